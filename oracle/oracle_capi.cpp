// oracle_capi.cpp — C entry points of the CPU oracle, for ctypes.
//
// TEST INFRASTRUCTURE ONLY.  Builds N independent oracle worlds from the same
// newt_options / newt_link_desc / newt_geom_desc records the product's C ABI
// takes (include/newt_gpu.h), steps them on host threads, and returns state,
// qacc, pairs and contacts for comparison with the HIP path.  Two
// instantiations: precision 0 = double (the FP64 oracle), 1 = float (newt's
// own arithmetic).
#include "flop_count.hpp"
#include "newt_world.hpp"

#include <atomic>
#include <chrono>
#include <memory>
#include <thread>

#include "../include/newt_gpu.h"

using namespace oracle;

namespace {

struct Base {
    virtual ~Base() = default;
    virtual void add_tree(const newt_link_desc* l, int n, int collide) = 0;
    virtual void add_geom(const newt_geom_desc* g) = 0;
    virtual int add_mesh(const double* v, int nv, const int* f, int nf) = 0;
    virtual void add_body(const newt_body_desc* b) = 0;
    virtual void finalize(int n_worlds) = 0;
    virtual int nq() const = 0;
    virtual int nv() const = 0;
    virtual void set_state(const double* qpos, const double* qvel) = 0;
    virtual void get_state(double* qpos, double* qvel) const = 0;
    virtual void get_qacc(double* qacc) const = 0;
    virtual void set_qfrc(const double* qfrc) = 0;
    virtual void get_qfrc(double* qfrc) const = 0;
    virtual void step(int n_steps, int threads) = 0;
    virtual int contacts(int world, int cap, int* ids, double* geo) const = 0;
    virtual int pairs(int world, int cap, int* out) const = 0;
    virtual int detect(int world) const = 0;
    virtual void mass_matrix(int world, double* out) const = 0;
    virtual void dynamics(int world, double* qacc, double* M, double* bias) const = 0;
    virtual int broadphase(int world, double* fat, double* bound) const = 0;
    virtual long long reinserts(int world, int reset) = 0;
    virtual void add_sensor(const newt_sensor_desc* d) = 0;
    virtual void set_magnetic(const double* f) = 0;
    virtual int sensor_dim() const = 0;
    virtual void get_sensordata(double* out) const = 0;
};

template <class T> struct Impl : Base {
    World<T> proto;
    std::vector<World<T>> worlds;

    explicit Impl(const newt_options& o) {
        proto.dt = T(o.dt);
        proto.gravity = v3(T(o.gravity[0]), T(o.gravity[1]), T(o.gravity[2]));
        proto.penetration_slop = T(o.penetration_slop);
        proto.integrator = o.integrator;
        proto.mode = o.solver_mode;
        proto.iterations = (uint32_t)o.iterations;
        proto.cone = o.cone;
    }
    void add_tree(const newt_link_desc* ld, int n, int collide) override {
        Tree<T> t;
        for (int i = 0; i < n; ++i) {
            const newt_link_desc& d = ld[i];
            Link<T> l;
            l.parent = d.parent;
            l.joint = d.joint;
            l.axis = v3(T(d.axis[0]), T(d.axis[1]), T(d.axis[2]));
            l.damping = T(d.damping);
            l.armature = T(d.armature);
            l.free_damping = T(d.free_damping);
            l.has_range = d.has_range != 0;
            l.range_lo = T(d.range[0]);
            l.range_hi = T(d.range[1]);
            // SolRef::DEFAULT (geom.rs:99-102), SolImp::DEFAULT (solver.rs:187-193)
            const double sr[2] = {0.02, 1.0}, si[4] = {0.9, 0.95, 0.001, 0.5};
            for (int k = 0; k < 2; ++k) l.limit_solref[k] = T(d.limit_custom ? d.limit_solref[k] : sr[k]);
            for (int k = 0; k < 4; ++k) l.limit_solimp[k] = T(d.limit_custom ? d.limit_solimp[k] : si[k]);
            l.limit_power = d.limit_custom ? d.limit_solimp_power : 2;
            // JointLimit::DEFAULT (joint.rs:152-157)
            l.limit_stiffness = T(d.limit_custom ? d.limit_stiffness : 1000.0);
            l.limit_damping = T(d.limit_custom ? d.limit_damping : 60.0);
            l.pj_pos = v3(T(d.parent_pos[0]), T(d.parent_pos[1]), T(d.parent_pos[2]));
            l.pj_quat = Q<T>{T(d.parent_quat[0]), T(d.parent_quat[1]), T(d.parent_quat[2]), T(d.parent_quat[3])};
            l.cj_pos = v3(T(d.child_pos[0]), T(d.child_pos[1]), T(d.child_pos[2]));
            l.cj_quat = Q<T>{T(d.child_quat[0]), T(d.child_quat[1]), T(d.child_quat[2]), T(d.child_quat[3])};
            l.mass = T(d.mass);
            for (int k = 0; k < 9; ++k) l.inertia.d[k] = T(d.inertia[k]);
            t.push_link(l);
        }
        proto.add_tree(t, collide != 0);
    }
    void add_geom(const newt_geom_desc* d) override {
        Geom<T> g;
        g.shape = d->shape;
        g.tree = d->tree;
        g.link = d->link;
        g.size = v3(T(d->size[0]), T(d->size[1]), T(d->size[2]));
        g.pos = v3(T(d->pos[0]), T(d->pos[1]), T(d->pos[2]));
        g.quat = Q<T>{T(d->quat[0]), T(d->quat[1]), T(d->quat[2]), T(d->quat[3])};
        g.friction = T(d->friction);
        g.margin = T(d->margin);
        g.gap = T(d->gap);
        g.solref = {T(d->solref[0]), T(d->solref[1])};
        g.solimp.dmin = T(d->solimp[0]);
        g.solimp.dmax = T(d->solimp[1]);
        g.solimp.width = T(d->solimp[2]);
        g.solimp.midpoint = T(d->solimp[3]);
        g.solimp.power = (uint32_t)d->solimp_power;
        g.condim = d->condim;
        g.group = d->group;
        g.mask = d->mask;
        g.body = d->body;
        g.mesh = d->shape == S_MESH ? d->mesh : -1;
        proto.add_geom(g);
    }
    int add_mesh(const double* v, int nv, const int* f, int nf) override {
        ConvexMesh<T> m;
        for (int i = 0; i < nv; ++i) m.v.push_back(v3(T(v[3 * i]), T(v[3 * i + 1]), T(v[3 * i + 2])));
        for (int i = 0; i < nf; ++i) m.f.push_back({f[3 * i], f[3 * i + 1], f[3 * i + 2]});
        return proto.add_mesh(m);
    }
    void add_body(const newt_body_desc* d) override {
        M3<T> inertia;
        for (int k = 0; k < 9; ++k) inertia.d[k] = T(d->inertia[k]);
        Body<T> b = make_body(T(d->mass), inertia, v3(T(d->position[0]), T(d->position[1]), T(d->position[2])),
                              Q<T>{T(d->orientation[0]), T(d->orientation[1]), T(d->orientation[2]), T(d->orientation[3])});
        b.v = v3(T(d->linear_velocity[0]), T(d->linear_velocity[1]), T(d->linear_velocity[2]));
        b.w = v3(T(d->angular_velocity_body[0]), T(d->angular_velocity_body[1]), T(d->angular_velocity_body[2]));
        b.has_max_lin = d->has_max_linear_velocity != 0;
        b.has_max_ang = d->has_max_angular_velocity != 0;
        b.max_lin = T(d->max_linear_velocity);
        b.max_ang = T(d->max_angular_velocity);
        b.gravity_scale = T(d->gravity_scale);
        proto.add_body(b);
    }
    void finalize(int n) override { worlds.assign(n, proto); }
    void add_sensor(const newt_sensor_desc* d) override {
        proto.add_sensor(d->kind, d->tree, d->link, d->body, d->geom, v3(T(d->offset[0]), T(d->offset[1]), T(d->offset[2])),
                         Q<T>{T(d->quat[0]), T(d->quat[1]), T(d->quat[2]), T(d->quat[3])});
    }
    void set_magnetic(const double* f) override { proto.magnetic_field = v3(T(f[0]), T(f[1]), T(f[2])); }
    int sensor_dim() const override { return (int)proto.sensordata.size(); }
    void get_sensordata(double* out) const override {
        const int n = sensor_dim();
        for (size_t w = 0; w < worlds.size(); ++w)
            for (int k = 0; k < n; ++k) out[w * n + k] = (double)worlds[w].sensordata[k];
    }
    int tree_nq() const {
        int s = 0;
        for (auto& t : proto.trees) s += t.nq();
        return s;
    }
    int tree_nv() const {
        int s = 0;
        for (auto& t : proto.trees) s += t.nv();
        return s;
    }
    int nq() const override { return tree_nq() + 7 * (int)proto.bodies.size(); }
    int nv() const override { return tree_nv() + 6 * (int)proto.bodies.size(); }
    void set_state(const double* qpos, const double* qvel) override {
        int NQ = nq(), NV = nv();
        for (size_t w = 0; w < worlds.size(); ++w) {
            int oq = 0, ov = 0;
            for (auto& t : worlds[w].trees) {
                for (int k = 0; k < t.nq(); ++k) t.q[k] = T(qpos[w * NQ + oq + k]);
                for (int k = 0; k < t.nv(); ++k) t.qdot[k] = T(qvel[w * NV + ov + k]);
                oq += t.nq();
                ov += t.nv();
            }
            for (auto& b : worlds[w].bodies) {
                const double* q = qpos + w * NQ + oq;
                const double* v = qvel + w * NV + ov;
                b.p = v3(T(q[0]), T(q[1]), T(q[2]));
                b.q = Q<T>{T(q[3]), T(q[4]), T(q[5]), T(q[6])};
                b.v = v3(T(v[0]), T(v[1]), T(v[2]));
                b.w = v3(T(v[3]), T(v[4]), T(v[5]));
                oq += 7;
                ov += 6;
            }
            // Public state setters refresh the query proxies (world.rs:1119-1136).
            worlds[w].refresh_broadphase();
        }
    }
    // Tree::qfrc_applied of every world, [world][nv].
    void set_qfrc(const double* qfrc) override {
        int NV = nv();
        for (size_t w = 0; w < worlds.size(); ++w) {
            int ov = 0;
            for (auto& t : worlds[w].trees) {
                for (int k = 0; k < t.nv(); ++k) t.qfrc_applied[k] = T(qfrc[w * NV + ov + k]);
                ov += t.nv();
            }
        }
    }
    void get_qfrc(double* qfrc) const override {
        int NV = nv();
        for (size_t w = 0; w < worlds.size(); ++w) {
            for (int k = 0; k < NV; ++k) qfrc[w * NV + k] = 0.0;
            int ov = 0;
            for (auto& t : worlds[w].trees) {
                for (int k = 0; k < t.nv(); ++k) qfrc[w * NV + ov + k] = (double)t.qfrc_applied[k];
                ov += t.nv();
            }
        }
    }
    void get_state(double* qpos, double* qvel) const override {
        int NQ = nq(), NV = nv();
        for (size_t w = 0; w < worlds.size(); ++w) {
            int oq = 0, ov = 0;
            for (auto& t : worlds[w].trees) {
                for (int k = 0; k < t.nq(); ++k) qpos[w * NQ + oq + k] = (double)t.q[k];
                for (int k = 0; k < t.nv(); ++k) qvel[w * NV + ov + k] = (double)t.qdot[k];
                oq += t.nq();
                ov += t.nv();
            }
            for (const auto& b : worlds[w].bodies) {
                double* q = qpos + w * NQ + oq;
                double* v = qvel + w * NV + ov;
                q[0] = (double)b.p.x; q[1] = (double)b.p.y; q[2] = (double)b.p.z;
                q[3] = (double)b.q.x; q[4] = (double)b.q.y; q[5] = (double)b.q.z; q[6] = (double)b.q.w;
                v[0] = (double)b.v.x; v[1] = (double)b.v.y; v[2] = (double)b.v.z;
                v[3] = (double)b.w.x; v[4] = (double)b.w.y; v[5] = (double)b.w.z;
                oq += 7;
                ov += 6;
            }
        }
    }
    void get_qacc(double* qacc) const override {
        int NV = nv();
        for (size_t w = 0; w < worlds.size(); ++w)
            for (int k = 0; k < NV; ++k)
                qacc[w * NV + k] = k < (int)worlds[w].last_qacc.size() ? (double)worlds[w].last_qacc[k] : 0.0;
    }
    void step(int n_steps, int threads) override {
        int n = (int)worlds.size();
        if (threads <= 1 || n <= 1) {
            for (int w = 0; w < n; ++w)
                for (int s = 0; s < n_steps; ++s) worlds[w].step();
            return;
        }
        std::atomic<int> next{0};
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k)
            pool.emplace_back([&]() {
                for (;;) {
                    int w = next.fetch_add(1);
                    if (w >= n) break;
                    for (int s = 0; s < n_steps; ++s) worlds[w].step();
                }
            });
        for (auto& th : pool) th.join();
    }
    // World::detect_contacts (world.rs:2055-2062) at the current state: a
    // fresh broadphase over a copy of the world, then the narrow phase.
    int detect(int world) const override {
        World<T> w = worlds[world];
        return (int)w.detect_contacts(w.active_pairs()).size();
    }
    int contacts(int world, int cap, int* ids, double* geo) const override {
        const auto& cs = worlds[world].last_contacts;
        int n = (int)cs.size();
        for (int k = 0; k < n && k < cap; ++k) {
            ids[3 * k] = cs[k].geom_a;
            ids[3 * k + 1] = cs[k].geom_b;
            ids[3 * k + 2] = cs[k].slot;
            double* g = geo + 7 * k;
            g[0] = (double)cs[k].pos.x; g[1] = (double)cs[k].pos.y; g[2] = (double)cs[k].pos.z;
            g[3] = (double)cs[k].normal.x; g[4] = (double)cs[k].normal.y; g[5] = (double)cs[k].normal.z;
            g[6] = (double)cs[k].penetration;
        }
        return n;
    }
    int pairs(int world, int cap, int* out) const override {
        const auto& ps = worlds[world].last_pairs;
        int n = (int)ps.size();
        for (int k = 0; k < n && k < cap; ++k) {
            out[2 * k] = ps[k].first;
            out[2 * k + 1] = ps[k].second;
        }
        return n;
    }
    // Explicit ABA (no external wrench), CRB mass matrix and RNE bias of
    // tree 0 at the world's current state.
    void dynamics(int world, double* qacc, double* Mout, double* bias) const override {
        const Tree<T>& t = worlds[world].trees[0];
        std::vector<Pose<T>> poses;
        forward_kinematics(t, poses);
        std::vector<std::pair<V3<T>, V3<T>>> ext(t.links.size(), {vzero<T>(), vzero<T>()});
        AbaWs<T> ws;
        std::vector<T> a;
        aba_tree(t, poses, worlds[world].gravity, ext, VI_EXPLICIT, worlds[world].dt, ws, a);
        std::vector<T> m = oracle::mass_matrix(t);
        std::vector<T> z(t.nv(), T(0));
        std::vector<T> h = inverse_dynamics(t, z, worlds[world].gravity, ext);
        for (int k = 0; k < t.nv(); ++k) { qacc[k] = (double)a[k]; bias[k] = (double)h[k]; }
        for (size_t k = 0; k < m.size(); ++k) Mout[k] = (double)m[k];
    }
    // Each geom's fat proxy bound and the swept bound it was last updated
    // with, [geom][min xyz, max xyz].
    int broadphase(int world, double* fat, double* bound) const override {
        const World<T>& w = worlds[world];
        for (size_t g = 0; g < w.fat.size(); ++g) {
            const Aabb<T>* src[2] = {&w.fat[g], &w.bound[g]};
            double* dst[2] = {fat + 6 * g, bound + 6 * g};
            for (int k = 0; k < 2; ++k) {
                dst[k][0] = (double)src[k]->mn.x; dst[k][1] = (double)src[k]->mn.y; dst[k][2] = (double)src[k]->mn.z;
                dst[k][3] = (double)src[k]->mx.x; dst[k][4] = (double)src[k]->mx.y; dst[k][5] = (double)src[k]->mx.z;
            }
        }
        return (int)w.fat.size();
    }
    long long reinserts(int world, int reset) override {
        const long long n = (long long)worlds[world].reinserts;
        if (reset) worlds[world].reinserts = 0;
        return n;
    }
    void mass_matrix(int world, double* out) const override {
        const Tree<T>& t = worlds[world].trees[0];
        std::vector<T> m = oracle::mass_matrix(t);
        for (size_t k = 0; k < m.size(); ++k) out[k] = (double)m[k];
    }
};

}  // namespace

// One narrow_phase call (dispatch_narrow_phase, contact.rs:2509-2590) on
// two geoms at the given world poses (x y z, qx qy qz qw), both mesh geoms
// naming the one mesh passed; contacts into out[k] = (position, normal,
// penetration).  precision 1 = float (newt's arithmetic), 0 = double.
namespace {
template <class T>
int narrow_one(const newt_geom_desc* da, const double* pa, const newt_geom_desc* db, const double* pb, const double* mv,
               int nv, const int* mf, int nf, int cap, double* out) {
    auto mesh = std::make_shared<ConvexMesh<T>>();
    for (int i = 0; i < nv; ++i) mesh->v.push_back(v3(T(mv[3 * i]), T(mv[3 * i + 1]), T(mv[3 * i + 2])));
    for (int i = 0; i < nf; ++i) mesh->f.push_back({mf[3 * i], mf[3 * i + 1], mf[3 * i + 2]});
    auto geom = [&](const newt_geom_desc* d) {
        Geom<T> g;
        g.shape = d->shape;
        g.size = v3(T(d->size[0]), T(d->size[1]), T(d->size[2]));
        g.friction = T(d->friction);
        g.margin = T(d->margin);
        g.gap = T(d->gap);
        if (g.shape == S_MESH) g.mesh_data = mesh;
        return g;
    };
    auto pose = [](const double* p) {
        return Pose<T>{v3(T(p[0]), T(p[1]), T(p[2])), Q<T>{T(p[3]), T(p[4]), T(p[5]), T(p[6])}};
    };
    std::vector<Contact<T>> buf;
    narrow_phase(0, geom(da), pose(pa), 1, geom(db), pose(pb), buf);
    int n = 0;
    for (const Contact<T>& c : buf) {
        if (n == cap) break;
        const double v[7] = {(double)c.pos.x, (double)c.pos.y, (double)c.pos.z, (double)c.normal.x,
                             (double)c.normal.y, (double)c.normal.z, (double)c.penetration};
        for (int k = 0; k < 7; ++k) out[7 * n + k] = v[k];
        ++n;
    }
    return n;
}
}  // namespace

extern "C" {

void* oracle_create(const newt_options* o, int precision) {
    if (precision == 1) return new Impl<float>(*o);
    if (precision == 2) return new Impl<Flop>(*o);
    return new Impl<double>(*o);
}
void oracle_destroy(void* h) { delete static_cast<Base*>(h); }
void oracle_add_tree(void* h, const newt_link_desc* l, int n, int collide) { static_cast<Base*>(h)->add_tree(l, n, collide); }
void oracle_add_geom(void* h, const newt_geom_desc* g) { static_cast<Base*>(h)->add_geom(g); }
void oracle_add_body(void* h, const newt_body_desc* b) { static_cast<Base*>(h)->add_body(b); }
int oracle_narrow(int precision, const newt_geom_desc* da, const double* pa, const newt_geom_desc* db, const double* pb,
                  const double* mv, int nv, const int* mf, int nf, int cap, double* out) {
    if (precision == 1) return narrow_one<float>(da, pa, db, pb, mv, nv, mf, nf, cap, out);
    return narrow_one<double>(da, pa, db, pb, mv, nv, mf, nf, cap, out);
}
int oracle_add_mesh(void* h, const double* v, int nv, const int* f, int nf) {
    return static_cast<Base*>(h)->add_mesh(v, nv, f, nf);
}
void oracle_finalize(void* h, int n_worlds) { static_cast<Base*>(h)->finalize(n_worlds); }
int oracle_nq(void* h) { return static_cast<Base*>(h)->nq(); }
int oracle_nv(void* h) { return static_cast<Base*>(h)->nv(); }
void oracle_set_state(void* h, const double* qpos, const double* qvel) { static_cast<Base*>(h)->set_state(qpos, qvel); }
void oracle_get_state(void* h, double* qpos, double* qvel) { static_cast<Base*>(h)->get_state(qpos, qvel); }
void oracle_get_qacc(void* h, double* qacc) { static_cast<Base*>(h)->get_qacc(qacc); }
void oracle_set_qfrc(void* h, const double* qfrc) { static_cast<Base*>(h)->set_qfrc(qfrc); }
void oracle_get_qfrc(void* h, double* qfrc) { static_cast<Base*>(h)->get_qfrc(qfrc); }
void oracle_step(void* h, int n_steps, int threads) { static_cast<Base*>(h)->step(n_steps, threads); }
int oracle_contacts(void* h, int world, int cap, int* ids, double* geo) { return static_cast<Base*>(h)->contacts(world, cap, ids, geo); }
int oracle_detect(void* h, int world) { return static_cast<Base*>(h)->detect(world); }
int oracle_pairs(void* h, int world, int cap, int* out) { return static_cast<Base*>(h)->pairs(world, cap, out); }
void oracle_mass_matrix(void* h, int world, double* out) { static_cast<Base*>(h)->mass_matrix(world, out); }
void oracle_dynamics(void* h, int world, double* qacc, double* M, double* bias) {
    static_cast<Base*>(h)->dynamics(world, qacc, M, bias);
}

// Operation counts of the calling thread's Flop arithmetic since the last
// reset: add/sub, mul, div, sqrt.
void oracle_flops(unsigned long long* out, int reset) {
    FlopCounts& c = flop_counts();
    out[0] = c.add;
    out[1] = c.mul;
    out[2] = c.div;
    out[3] = c.sqrt;
    if (reset) c = FlopCounts{};
}

// Deterministic sin/cos of the oracle, for the kernel's polynomial check.
int oracle_broadphase(void* h, int world, double* fat, double* bound) {
    return static_cast<Base*>(h)->broadphase(world, fat, bound);
}
long long oracle_reinserts(void* h, int world, int reset) { return static_cast<Base*>(h)->reinserts(world, reset); }
void oracle_add_sensor(void* h, const newt_sensor_desc* d) { static_cast<Base*>(h)->add_sensor(d); }
void oracle_set_magnetic(void* h, const double* f) { static_cast<Base*>(h)->set_magnetic(f); }
int oracle_sensor_dim(void* h) { return static_cast<Base*>(h)->sensor_dim(); }
void oracle_get_sensordata(void* h, double* out) { static_cast<Base*>(h)->get_sensordata(out); }
// The dynamic tree's pair set for standalone boxes (DynamicAabbTree::insert
// + compute_pairs, broadphase.rs:46-54, 384-455), in f32: every box is
// fattened at insertion and the pairs are the sorted (a, b), a < b, whose
// fat bounds overlap.  boxes: [n][min xyz, max xyz].
int oracle_box_pairs_f32(int n, const float* boxes, int cap, int* out) {
    std::vector<Aabb<float>> fat((size_t)n);
    for (int i = 0; i < n; ++i) {
        const float* b = boxes + 6 * i;
        fat[i] = fatten(Aabb<float>{{b[0], b[1], b[2]}, {b[3], b[4], b[5]}});
    }
    int k = 0;
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b)
            if (overlaps(fat[a], fat[b])) {
                if (k < cap) {
                    out[2 * k] = a;
                    out[2 * k + 1] = b;
                }
                ++k;
            }
    return k;
}
double oracle_nsin64(double x) { return nsin(x); }
double oracle_ncos64(double x) { return ncos(x); }
float oracle_nsin32(float x) { return nsin(x); }
float oracle_ncos32(float x) { return ncos(x); }
}
