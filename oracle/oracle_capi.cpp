// CPU ORACLE -- test infrastructure only (see pe_oracle.hpp).
// C entry points over the restatement, in FP32 (peo32_*) and FP64 (peo64_*).
#include "pe_oracle.hpp"
#include "../include/pe_amd.h"
#include <cstdio>

using namespace pe_oracle;

namespace {
template <class R> struct Handle {
    World<R> w;
    Solver<R> s;
};

template <class R> Handle<R>* create(const pe_scene_desc* sc, const pe_world_params* prm) {
    auto* h = new Handle<R>();
    World<R>& w = h->w;
    w.dt = R(prm->dt);
    w.gravity = V3<R>(R(prm->gravity[0]), R(prm->gravity[1]), R(prm->gravity[2]));
    w.iterations = prm->iterations;
    w.bodies.resize(sc->n);
    // one shape per (mesh, kind), shared by the bodies that name it
    std::vector<std::shared_ptr<Poly<R>>> convex(sc->n_meshes), concave(sc->n_meshes);
    auto mesh_poly = [&](int m, bool is_convex) {
        auto& slot = is_convex ? convex[m] : concave[m];
        if (!slot) {
            slot = std::make_shared<Poly<R>>();
            const pe_mesh_desc& md = sc->meshes[m];
            for (int v = 0; v < md.nv; v++)
                slot->v.push_back(V3<R>(R(md.vertices[3 * v]), R(md.vertices[3 * v + 1]), R(md.vertices[3 * v + 2])));
            for (int f = 0; f < md.nf; f++) {
                std::vector<int> idx(md.face_index + md.face_start[f], md.face_index + md.face_start[f + 1]);
                slot->f.push_back(idx);
                slot->fn.push_back(V3<R>(R(md.face_normal[3 * f]), R(md.face_normal[3 * f + 1]), R(md.face_normal[3 * f + 2])));
            }
            if (is_convex) convex_set_mesh(*slot);
            else concave_set_mesh(*slot);
            w.polys.push_back(slot);
        }
        return slot.get();
    };
    for (int i = 0; i < sc->n; i++) {
        Body<R>& b = w.bodies[i];
        b.type = sc->shape_type[i];
        const double* sp = sc->shape_param + 4 * i;
        if (b.type == ST_BOX) {
            for (int k = 0; k < 3; k++) { b.size[k] = R(sp[k]); b.half[k] = b.size[k] / R(2); }
            w.polys.push_back(std::make_shared<Poly<R>>(box_poly(b)));
            b.poly = w.polys.back().get();
        } else if (b.type == ST_SPHERE) {
            b.radius = R(sp[0]);
        } else if (b.type == ST_CYLINDER) {
            b.radius = R(sp[0]); b.height = R(sp[1]);
            w.polys.push_back(std::make_shared<Poly<R>>(cylinder_poly(b)));
            b.poly = w.polys.back().get();
        } else if (b.type == ST_COMPOUND) {
            // CompoundShape::addShape for each child, then CompoundShape::init
            for (int k = sc->compound_start[i]; k < sc->compound_start[i + 1]; k++) {
                const pe_compound_child& ch = sc->compound_children[k];
                Body<R> kid;
                kid.type = ch.shape_type;
                if (kid.type == ST_BOX) {
                    for (int q = 0; q < 3; q++) { kid.size[q] = R(ch.shape_param[q]); kid.half[q] = kid.size[q] / R(2); }
                    w.polys.push_back(std::make_shared<Poly<R>>(box_poly(kid)));
                    kid.poly = w.polys.back().get();
                } else if (kid.type == ST_SPHERE) {
                    kid.radius = R(ch.shape_param[0]);
                } else if (kid.type == ST_CYLINDER) {
                    kid.radius = R(ch.shape_param[0]); kid.height = R(ch.shape_param[1]);
                    w.polys.push_back(std::make_shared<Poly<R>>(cylinder_poly(kid)));
                    kid.poly = w.polys.back().get();
                } else {
                    kid.poly = mesh_poly(ch.mesh_index, true);
                }
                for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) kid.xf.basis(r, c) = R(ch.rotation[3 * r + c]);
                for (int q = 0; q < 3; q++) kid.xf.origin[q] = R(ch.position[q]);
                b.kids.push_back(kid);
                b.ratios.push_back(R(ch.mass_ratio));
            }
            compound_init(b);
        } else {
            b.poly = mesh_poly(sc->mesh_index[i], b.type == ST_CONVEX);
        }
        for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) b.xf.basis(r, c) = R(sc->rotation[9 * i + 3 * r + c]);
        for (int k = 0; k < 3; k++) {
            b.xf.origin[k] = R(sc->position[3 * i + k]);
            b.v[k] = R(sc->linear_velocity[3 * i + k]);
            b.w[k] = R(sc->angular_velocity[3 * i + k]);
        }
        body_set_mass_props(b, R(sc->mass[i]));
        body_update_world_inertia(b);
        b.mu = R(sc->friction[i]);
        b.e = R(sc->restitution[i]);
        b.lin_damp = sc->linear_damping ? R(sc->linear_damping[i]) : R(0);
        b.ang_damp = sc->angular_damping ? R(sc->angular_damping[i]) : R(0);
        b.kinematic = sc->kinematic[i] != 0;
        b.ignore = sc->ignore_collision && sc->ignore_collision[i] != 0;
        if (sc->lifetime) b.life_time = R(sc->lifetime[i]);
    }
    return h;
}

template <class R> void get_state(Handle<R>* h, double* pos, double* rot, double* v, double* w) {
    for (size_t i = 0; i < h->w.bodies.size(); i++) {
        const Body<R>& b = h->w.bodies[i];
        for (int k = 0; k < 3; k++) {
            if (pos) pos[3 * i + k] = b.xf.origin[k];
            if (v) v[3 * i + k] = b.v[k];
            if (w) w[3 * i + k] = b.w[k];
        }
        if (rot) for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) rot[9 * i + 3 * r + c] = b.xf.basis(r, c);
    }
}
template <class R> void set_state(Handle<R>* h, const double* pos, const double* rot, const double* v, const double* w) {
    for (size_t i = 0; i < h->w.bodies.size(); i++) {
        Body<R>& b = h->w.bodies[i];
        for (int k = 0; k < 3; k++) {
            if (pos) b.xf.origin[k] = R(pos[3 * i + k]);
            if (v) b.v[k] = R(v[3 * i + k]);
            if (w) b.w[k] = R(w[3 * i + k]);
        }
        if (rot) for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) b.xf.basis(r, c) = R(rot[9 * i + 3 * r + c]);
        body_update_world_inertia(b);
    }
}
template <class R> void get_contacts(Handle<R>* h, int32_t* npts, int32_t* swap, double* pts) {
    const auto& res = h->w.results;
    for (size_t i = 0; i < res.size(); i++) {
        if (npts) npts[i] = res[i].npts;
        if (swap) swap[i] = res[i].swap;
        if (!pts) continue;
        for (int k = 0; k < 8; k++) {
            const ContactPoint<R>& cp = res[i].pts[k];
            double* o = pts + (i * 8 + k) * 13;
            for (int c = 0; c < 3; c++) {
                o[c] = cp.wpos[c]; o[3 + c] = cp.normal[c]; o[6 + c] = cp.la[c]; o[9 + c] = cp.lb[c];
            }
            o[12] = cp.dist;
        }
    }
}
// Polyhedral data of a body after shape set-up: counts[4] = vertices, faces, face indices,
// unique edges; then vertex positions, face starts / indices / normals, edges (6 per edge) and
// volume + unit-mass inertia (10).  Any output pointer may be null.
template <class R> int get_poly(Handle<R>* h, int body, int32_t* counts, double* v, int32_t* fstart, int32_t* fidx,
                                double* fn, double* e, double* vi) {
    const Poly<R>* p = h->w.bodies[body].poly;
    if (!p) return -1;
    int nidx = 0;
    for (auto& f : p->f) nidx += (int)f.size();
    if (counts) { counts[0] = (int)p->v.size(); counts[1] = (int)p->f.size(); counts[2] = nidx; counts[3] = (int)p->e.size(); }
    if (v) for (size_t i = 0; i < p->v.size(); i++) for (int k = 0; k < 3; k++) v[3 * i + k] = p->v[i][k];
    int o = 0;
    for (size_t f = 0; f < p->f.size(); f++) {
        if (fstart) fstart[f] = o;
        for (int x : p->f[f]) { if (fidx) fidx[o] = x; o++; }
        if (fn) for (int k = 0; k < 3; k++) fn[3 * f + k] = p->fn[f][k];
    }
    if (fstart) fstart[p->f.size()] = o;
    if (e) for (size_t i = 0; i < p->e.size(); i++) for (int k = 0; k < 3; k++) { e[6 * i + k] = p->e[i].first[k]; e[6 * i + 3 + k] = p->e[i].second[k]; }
    if (vi) { vi[0] = p->volume; for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) vi[1 + 3 * r + c] = p->inertia(r, c); }
    return 0;
}
}  // namespace

#define PEO_API(P, R)                                                                                     \
    extern "C" void* P##_world_create(const pe_scene_desc* sc, const pe_world_params* prm) {             \
        return create<R>(sc, prm);                                                                       \
    }                                                                                                    \
    extern "C" void P##_world_destroy(void* h) { delete (Handle<R>*)h; }                                \
    extern "C" void P##_add_joint(void* h, const pe_joint_desc* d) {                                     \
        auto& w = ((Handle<R>*)h)->w;                                                                    \
        Joint<R> j;                                                                                      \
        j.type = d->type; j.a = d->body_a; j.b = d->body_b;                                             \
        for (int k = 0; k < 3; k++) {                                                                    \
            j.anchor_a[k] = R(d->anchor_a[k]); j.anchor_b[k] = R(d->anchor_b[k]);                       \
            j.axis_a[k] = R(d->axis_a[k]); j.axis_b[k] = R(d->axis_b[k]);                               \
        }                                                                                                \
        j.use_limits = d->use_limits != 0; j.use_motor = d->use_motor != 0;                             \
        j.min_angle = R(d->min_angle); j.max_angle = R(d->max_angle); j.target_speed = R(d->target_speed); \
        w.joints.push_back(j);                                                                           \
    }                                                                                                    \
    extern "C" void P##_add_ignore(void* h, int a, int b) {                                              \
        ((Handle<R>*)h)->w.ignore.insert(std::make_pair(std::min(a, b), std::max(a, b)));               \
    }                                                                                                    \
    extern "C" int P##_get_joint_order(void* h, int32_t* out) {                                          \
        auto& w = ((Handle<R>*)h)->w;                                                                    \
        for (size_t k = 0; k < w.joint_order.size(); k++) out[k] = w.joint_order[k];                    \
        return (int)w.joint_order.size();                                                                \
    }                                                                                                    \
    extern "C" void P##_set_color_order(void* h, int on) { ((Handle<R>*)h)->w.use_color_order = on != 0; } \
    extern "C" void P##_set_sleep(void* h, double lin2, double ang2, double t) {                         \
        auto& w = ((Handle<R>*)h)->w;                                                                    \
        w.sleep_lin2 = R(lin2); w.sleep_ang2 = R(ang2); w.sleep_time = R(t);                             \
    }                                                                                                    \
    extern "C" void P##_get_status(void* h, int32_t* status, int32_t* scount, int32_t* dcount, double* st) { \
        auto& w = ((Handle<R>*)h)->w;                                                                    \
        for (size_t i = 0; i < w.bodies.size(); i++) {                                                   \
            const Body<R>& b = w.bodies[i];                                                              \
            if (status) status[i] = (b.sleep ? 1 : 0) | (b.removed ? 2 : 0);                             \
            if (scount) scount[i] = b.static_count;                                                      \
            if (dcount) dcount[i] = b.dynamic_count;                                                     \
            if (st) st[i] = b.sleep_time;                                                                \
        }                                                                                                \
    }                                                                                                    \
    extern "C" void P##_set_report(void* h, int body, int on) { ((Handle<R>*)h)->w.bodies[body].report = on != 0; } \
    extern "C" void P##_remove_body(void* h, int body) { ((Handle<R>*)h)->w.bodies[body].removed = true; } \
    extern "C" int P##_raycast(void* h, const double* ray, int nignore, const int32_t* ignore, double* out) { \
        std::set<int> ig(ignore, ignore + nignore);                                                      \
        const RayHit<R> r = raycast(((Handle<R>*)h)->w, V3<R>(R(ray[0]), R(ray[1]), R(ray[2])),          \
                                    V3<R>(R(ray[3]), R(ray[4]), R(ray[5])), R(ray[6]), ig);              \
        out[0] = r.body; out[1] = r.distance;                                                            \
        for (int k = 0; k < 3; k++) { out[2 + k] = r.point[k]; out[5 + k] = r.normal[k]; }                \
        return r.hit ? 1 : 0;                                                                            \
    }                                                                                                    \
    extern "C" int P##_num_events(void* h) { return (int)((Handle<R>*)h)->w.events.size(); }            \
    extern "C" void P##_get_events(void* h, int32_t* ab, double* data) {                                 \
        auto& ev = ((Handle<R>*)h)->w.events;                                                            \
        for (size_t i = 0; i < ev.size(); i++) {                                                         \
            ab[2 * i] = ev[i].a; ab[2 * i + 1] = ev[i].b;                                                \
            double* o = data + 10 * i;                                                                   \
            for (int k = 0; k < 3; k++) { o[k] = ev[i].pos[k]; o[3 + k] = ev[i].nor[k]; o[7 + k] = ev[i].vel[k]; } \
            o[6] = ev[i].depth;                                                                          \
        }                                                                                                \
    }                                                                                                    \
    extern "C" void P##_step(void* h, int nsteps) {                                                      \
        auto* hh = (Handle<R>*)h;                                                                        \
        for (int i = 0; i < nsteps; i++) world_step(hh->w, hh->s);                                       \
    }                                                                                                    \
    extern "C" void P##_integrate(void* h) { stage_integrate(((Handle<R>*)h)->w); }                     \
    extern "C" int P##_broadphase(void* h) {                                                             \
        stage_broadphase(((Handle<R>*)h)->w);                                                            \
        return (int)((Handle<R>*)h)->w.pairs.size();                                                     \
    }                                                                                                    \
    extern "C" void P##_narrowphase(void* h) { stage_narrowphase(((Handle<R>*)h)->w); }                 \
    extern "C" void P##_solve(void* h) {                                                                 \
        auto* hh = (Handle<R>*)h;                                                                        \
        stage_solver_setup(hh->w, hh->s);                                                                \
        if (hh->w.use_color_order) compute_color_order(hh->w);                                           \
        stage_solve(hh->w, hh->s);                                                                       \
    }                                                                                                    \
    extern "C" int P##_num_bodies(void* h) { return (int)((Handle<R>*)h)->w.bodies.size(); }            \
    extern "C" int P##_num_pairs(void* h) { return (int)((Handle<R>*)h)->w.pairs.size(); }              \
    extern "C" int P##_get_axis(void* h) { return ((Handle<R>*)h)->w.axis; }                            \
    extern "C" void P##_set_axis(void* h, int a) { ((Handle<R>*)h)->w.axis = a; }                       \
    extern "C" void P##_get_pairs(void* h, int32_t* ab) {                                                \
        auto& p = ((Handle<R>*)h)->w.pairs;                                                              \
        for (size_t i = 0; i < p.size(); i++) { ab[2 * i] = p[i].first; ab[2 * i + 1] = p[i].second; }   \
    }                                                                                                    \
    extern "C" void P##_set_pairs(void* h, int n, const int32_t* ab) {                                   \
        auto& p = ((Handle<R>*)h)->w.pairs;                                                              \
        p.resize(n);                                                                                     \
        for (int i = 0; i < n; i++) p[i] = {ab[2 * i], ab[2 * i + 1]};                                   \
    }                                                                                                    \
    extern "C" void P##_get_state(void* h, double* pos, double* rot, double* v, double* w) {             \
        get_state((Handle<R>*)h, pos, rot, v, w);                                                        \
    }                                                                                                    \
    extern "C" void P##_set_state(void* h, const double* pos, const double* rot, const double* v,        \
                                  const double* w) {                                                     \
        set_state((Handle<R>*)h, pos, rot, v, w);                                                        \
    }                                                                                                    \
    extern "C" void P##_get_aabbs(void* h, double* mn, double* mx) {                                     \
        auto& bs = ((Handle<R>*)h)->w.bodies;                                                            \
        for (size_t i = 0; i < bs.size(); i++)                                                           \
            for (int k = 0; k < 3; k++) { mn[3 * i + k] = bs[i].aabb_min[k]; mx[3 * i + k] = bs[i].aabb_max[k]; } \
    }                                                                                                    \
    extern "C" void P##_get_world_inv_inertia(void* h, double* out) {                                    \
        auto& bs = ((Handle<R>*)h)->w.bodies;                                                            \
        for (size_t i = 0; i < bs.size(); i++)                                                           \
            for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) out[9 * i + 3 * r + c] = bs[i].world_inv_inertia(r, c); \
    }                                                                                                    \
    extern "C" void P##_get_mass_props(void* h, double* inv_mass, double* local_inv_inertia) {           \
        auto& bs = ((Handle<R>*)h)->w.bodies;                                                            \
        for (size_t i = 0; i < bs.size(); i++) {                                                         \
            inv_mass[i] = bs[i].inv_mass;                                                                \
            for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++)                                      \
                local_inv_inertia[9 * i + 3 * r + c] = bs[i].local_inv_inertia(r, c);                    \
        }                                                                                                \
    }                                                                                                    \
    extern "C" int P##_num_results(void* h) { return (int)((Handle<R>*)h)->w.results.size(); }          \
    extern "C" void P##_get_contacts(void* h, int32_t* npts, int32_t* swap, double* pts) {               \
        get_contacts((Handle<R>*)h, npts, swap, pts);                                                    \
    }                                                                                                    \
    extern "C" int P##_get_poly(void* h, int body, int32_t* counts, double* v, int32_t* fstart, int32_t* fidx, \
                                double* fn, double* e, double* vi) {                                     \
        return get_poly((Handle<R>*)h, body, counts, v, fstart, fidx, fn, e, vi);                        \
    }                                                                                                    \
    extern "C" int P##_num_colors(void* h) { return ((Handle<R>*)h)->w.ncolors; }                       \
    extern "C" int P##_get_solve_order(void* h, int32_t* order) {                                        \
        auto& o = ((Handle<R>*)h)->w.solve_order;                                                        \
        for (size_t i = 0; i < o.size(); i++) order[i] = o[i];                                           \
        return (int)o.size();                                                                            \
    }                                                                                                    \
    extern "C" void P##_get_color_of(void* h, int32_t* col) {                                            \
        auto& c = ((Handle<R>*)h)->w.color_of;                                                           \
        for (size_t i = 0; i < c.size(); i++) col[i] = c[i];                                             \
    }                                                                                                    \
    extern "C" void P##_get_impulses(void* h, double* out) { /* per result: 8 x (ln, lt0, lt1) */       \
        auto* hh = (Handle<R>*)h;                                                                        \
        for (size_t i = 0; i < hh->s.rows.size(); i++)                                                   \
            for (size_t k = 0; k < 8; k++) {                                                             \
                double* o = out + (i * 8 + k) * 3;                                                       \
                if (k < hh->s.rows[i].size()) {                                                          \
                    o[0] = hh->s.rows[i][k].ln; o[1] = hh->s.rows[i][k].lt0; o[2] = hh->s.rows[i][k].lt1; \
                } else { o[0] = o[1] = o[2] = 0; }                                                       \
            }                                                                                            \
    }

PEO_API(peo32, float)
PEO_API(peo64, double)
