// Host/device shared definitions: SoA field layout of the body store, error handling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace pe {

// Body store: one float array of NF fields x stride (stride = n rounded up to 64), field-major,
// so every per-field access of consecutive bodies is one coalesced load.
enum BodyField : int {
    F_POS = 0,     // 3: origin
    F_ROT = 3,     // 9: basis, row-major
    F_V = 12,      // 3: linear velocity
    F_W = 15,      // 3: angular velocity
    F_IW = 18,     // 9: world inverse inertia
    F_AMIN = 27,   // 3: AABB min
    F_AMAX = 30,   // 3: AABB max
    F_IM = 33,     // inverse mass
    F_MASS = 34,   // mass
    F_LINV = 35,   // 9: local inverse inertia
    F_SP = 44,     // 4: shape parameters (box size / sphere radius / cylinder radius, height)
    F_MU = 48,     // friction
    F_E = 49,      // restitution
    F_LD = 50,     // linear damping
    F_AD = 51,     // angular damping
    F_TV = 52,     // 3: temp linear velocity (solver)
    F_TW = 55,     // 3: temp angular velocity (solver)
    F_SCALE = 58,  // AABB scale of the current pose
    NF = 59
};
enum BodyIntField : int { I_TYPE = 0, I_KIN = 1, NI = 2 };

// Contact slot fields: one float array of CF fields x (pair capacity * 8)
enum ContactField : int { C_WPOS = 0, C_N = 3, C_LA = 6, C_LB = 9, C_D = 12, CF = 13 };
// Solver row fields: one float array of RF fields x (pair capacity * 8)
enum RowField : int {
    R_RA = 0, R_RB = 3, R_N = 6, R_T0 = 9, R_T1 = 12, R_RHS = 15, R_MN = 16, R_MT0 = 17, R_MT1 = 18,
    R_MU = 19, R_LN = 20, R_LT0 = 21, R_LT1 = 22, RF = 23
};

struct BodyView {
    float* f;
    int32_t* i;
    int n;
    int stride;
    __host__ __device__ float& at(int field, int b) const { return f[(size_t)field * stride + b]; }
    __host__ __device__ int32_t& iat(int field, int b) const { return i[(size_t)field * stride + b]; }
};

}  // namespace pe

// ------------------------------------------------------------------ host error helpers
void pe_set_error(const std::string& msg);
#define PE_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            pe_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return PE_ERR_DEVICE;                                                        \
        }                                                                                \
    } while (0)
