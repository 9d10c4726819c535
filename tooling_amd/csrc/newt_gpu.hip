// newt_gpu.hip — host model compiler, device state and the C ABI of
// include/newt_gpu.h, driving the lane-per-world step kernels for gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/newt_gpu.h"
#include "newt_step.hpp"

using namespace newt;

namespace {

thread_local std::string g_err;

newt_status fail(newt_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

#define HIP_TRY(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(NEWT_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 64;  // one wavefront: 64 worlds

// Largest contact count one candidate pair can emit (ContactBuf, contact.rs:108-124).
int pair_max_contacts(int sa, int sb) {
    int p = sa == SH_PLANE ? sb : sa;
    if (p == SH_SPHERE) return 1;
    if (p == SH_CAPSULE) return 2;
    return 4;
}

}  // namespace

struct newt_model {
    newt_options opt{};
    std::vector<std::vector<newt_link_desc>> trees;
    std::vector<int> tree_collide;
    std::vector<newt_geom_desc> geoms;
    bool finalized = false;
    // compiled
    ModelT dev{};
    std::vector<LinkT> links;
    std::vector<GeomT> hgeoms;
    std::vector<PairT> pairs;
    std::vector<int> chain;
    LinkT* d_links = nullptr;
    GeomT* d_geoms = nullptr;
    PairT* d_pairs = nullptr;
    int* d_chain = nullptr;
    std::vector<double> q_default;
};

struct newt_data {
    const newt_model* model = nullptr;
    int W = 0;
    int precision = 0;
    Layout L{};
    DataT d{};
    void* ws = nullptr;      // real workspace (double or float)
    double* dws = nullptr;   // factorization workspace
    int* iws = nullptr;      // integer workspace
    hipStream_t last_stream = nullptr;
};

namespace {

// Parent->child transform of a fixed joint (tree.rs:1124-1131), host double.
void fixed_xup(const newt_link_desc& d, double E[9], double r[3]) {
    Qt<double> pq{d.parent_quat[0], d.parent_quat[1], d.parent_quat[2], d.parent_quat[3]};
    Qt<double> cq{d.child_quat[0], d.child_quat[1], d.child_quat[2], d.child_quat[3]};
    Qt<double> rel = qmul(pq, qconj(cq));
    M33<double> m = qmat(qconj(rel));
    for (int k = 0; k < 9; ++k) E[k] = m.m[k];
    V3<double> pj = mk3(d.parent_pos[0], d.parent_pos[1], d.parent_pos[2]);
    V3<double> cj = mk3(d.child_pos[0], d.child_pos[1], d.child_pos[2]);
    V3<double> t = cj - mv(m, pj);
    r[0] = t.x; r[1] = t.y; r[2] = t.z;
}

// Mat6::from_spatial_inertia with a zero COM offset (spatial.rs:355-381),
// evaluated with the reference's operations so the signs of its zero
// entries match: [I + ([c]x[c]x)(-m), [c]x m; [c]x (-m), m 1], c = 0.
void spatial_inertia(const newt_link_desc& d, double si[36]) {
    const double m = d.mass;
    const double z = 0.0;
    // skew(0) in column-major order (math.rs:418-424)
    const double cx[9] = {0.0, z, -z, -z, 0.0, z, z, -z, 0.0};
    auto get = [&](const double* a, int r, int c) { return a[c * 3 + r]; };
    double cxcx[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += get(cx, r, k) * get(cx, k, c);
            cxcx[c * 3 + r] = s;
        }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            si[r * 6 + c] = d.inertia[c * 3 + r] + get(cxcx, r, c) * (-m);
            si[r * 6 + c + 3] = get(cx, r, c) * m;
            si[(r + 3) * 6 + c] = get(cx, r, c) * (-m);
            si[(r + 3) * 6 + c + 3] = r == c ? m : 0.0;
        }
}

newt_status compile(newt_model* M) {
    if (M->trees.size() != 1)
        return fail(NEWT_ERR_UNSUPPORTED, "this build steps models with exactly one tree");
    const newt_options& o = M->opt;
    if (o.iterations < 1) return fail(NEWT_ERR_INVALID_CONFIG, "solver iterations must be >= 1");
    if (o.solver_mode == NEWT_SOLVER_NEWTON && o.cone == NEWT_CONE_ELLIPTIC)
        return fail(NEWT_ERR_INVALID_CONFIG, "solver=newton with cone=elliptic is not supported yet; use cone=pyramidal");
    if (o.integrator != NEWT_INTEGRATOR_EULER)
        return fail(NEWT_ERR_UNSUPPORTED, "the GPU step integrates with semi-implicit Euler only");
    if (o.solver_mode == NEWT_SOLVER_NEWTON) return fail(NEWT_ERR_UNSUPPORTED, "Newton solver is not built for trees");
    if (o.cone != NEWT_CONE_PYRAMIDAL) return fail(NEWT_ERR_UNSUPPORTED, "elliptic cone is not built");
    const auto& T = M->trees[0];
    int nq = 0, nv = 0;
    M->links.clear();
    M->chain.clear();
    M->q_default.clear();
    for (size_t i = 0; i < T.size(); ++i) {
        const newt_link_desc& d = T[i];
        if (i == 0 && !(d.joint == NEWT_JOINT_FREE || d.joint == NEWT_JOINT_FIXED))
            return fail(NEWT_ERR_INVALID_ARG, "root joint must be Free or Fixed");
        if (i > 0 && !(d.parent >= 0 && d.parent < (int)i))
            return fail(NEWT_ERR_INVALID_ARG, "parent index must precede its link");
        if (i > 0 && d.joint == NEWT_JOINT_FREE) return fail(NEWT_ERR_INVALID_ARG, "free joint only at the root");
        if (d.joint == NEWT_JOINT_BALL) return fail(NEWT_ERR_UNSUPPORTED, "ball joints are not built on the GPU path");
        if (d.has_range) return fail(NEWT_ERR_UNSUPPORTED, "joint ranges (limit rows) are not built on the GPU path");
        LinkT l{};
        l.parent = d.parent;
        l.joint = d.joint;
        l.q_off = nq;
        l.v_off = nv;
        for (int k = 0; k < 3; ++k) {
            l.axis[k] = d.axis[k];
            l.pj[k] = d.parent_pos[k];
            l.cj[k] = d.child_pos[k];
        }
        for (int k = 0; k < 4; ++k) {
            l.pjq[k] = d.parent_quat[k];
            l.cjq[k] = d.child_quat[k];
        }
        fixed_xup(d, l.xfE, l.xfr);
        spatial_inertia(d, l.si);
        l.mass = d.mass;
        l.damping = d.damping;
        l.armature = d.armature;
        l.free_damping = d.free_damping;
        // ancestor chain root..self
        std::vector<int> ch;
        for (int j = (int)i; j >= 0; j = T[j].parent) ch.push_back(j);
        std::reverse(ch.begin(), ch.end());
        l.chain_off = (int)M->chain.size();
        l.chain_len = (int)ch.size();
        M->chain.insert(M->chain.end(), ch.begin(), ch.end());
        // default q (tree.rs:275-330)
        if (d.joint == NEWT_JOINT_FREE) {
            for (int k = 0; k < 3; ++k) M->q_default.push_back(d.parent_pos[k]);
            for (int k = 0; k < 4; ++k) M->q_default.push_back(d.parent_quat[k]);
            nq += 7;
            nv += 6;
        } else if (d.joint == NEWT_JOINT_HINGE || d.joint == NEWT_JOINT_SLIDE) {
            M->q_default.push_back(0.0);
            nq += 1;
            nv += 1;
        }
        M->links.push_back(l);
    }
    // geoms
    M->hgeoms.clear();
    for (const newt_geom_desc& g : M->geoms) {
        if (g.shape < NEWT_SHAPE_PLANE || g.shape > NEWT_SHAPE_CAPSULE) return fail(NEWT_ERR_INVALID_ARG, "unknown geom shape");
        if (g.tree >= 1) return fail(NEWT_ERR_INVALID_ARG, "geom attached to a missing tree");
        if (g.tree == 0 && (g.link < 0 || g.link >= (int)T.size())) return fail(NEWT_ERR_INVALID_ARG, "geom link out of range");
        if (g.condim != 3 && g.condim != 1 && g.condim != 4 && g.condim != 6)
            return fail(NEWT_ERR_INVALID_ARG, "condim must be 1, 3, 4 or 6");
        GeomT h{};
        h.shape = g.shape;
        h.link = g.tree == 0 ? g.link : -1;
        h.condim = g.condim;
        for (int k = 0; k < 3; ++k) {
            h.size[k] = g.size[k];
            h.pos[k] = g.pos[k];
        }
        for (int k = 0; k < 4; ++k) h.quat[k] = g.quat[k];
        h.friction = g.friction;
        h.margin = g.margin;
        h.gap = g.gap;
        h.solref[0] = g.solref[0];
        h.solref[1] = g.solref[1];
        for (int k = 0; k < 4; ++k) h.solimp[k] = g.solimp[k];
        h.power = g.solimp_power;
        M->hgeoms.push_back(h);
    }
    // Static candidate pairs: group/mask filter, disabled self collision,
    // same attachment (broadphase.rs:16-18, 426-437; world.rs:1219-1225).
    M->pairs.clear();
    int cap_c = 0;
    bool self_collide = M->tree_collide[0] != 0;
    for (size_t a = 0; a < M->geoms.size(); ++a)
        for (size_t b = a + 1; b < M->geoms.size(); ++b) {
            const newt_geom_desc& ga = M->geoms[a];
            const newt_geom_desc& gb = M->geoms[b];
            if (!((ga.group & gb.mask) != 0 && (gb.group & ga.mask) != 0)) continue;
            bool la = ga.tree >= 0, lb = gb.tree >= 0;
            if (la && lb && !self_collide) continue;
            if (la == lb && (!la || ga.link == gb.link)) continue;
            bool plane_pair = (ga.shape == NEWT_SHAPE_PLANE) != (gb.shape == NEWT_SHAPE_PLANE);
            if (!plane_pair)
                return fail(NEWT_ERR_UNSUPPORTED_PAIR,
                            "contact pair geom " + std::to_string(a) + " x geom " + std::to_string(b) +
                                " is not built on the GPU path (plane-primitive pairs only)");
            if (std::min(ga.condim, gb.condim) != 3)
                return fail(NEWT_ERR_UNSUPPORTED, "only condim 3 pyramidal contacts are built");
            M->pairs.push_back({(int)a, (int)b});
            cap_c += pair_max_contacts(ga.shape, gb.shape);
        }
    if (o.solver_mode == NEWT_SOLVER_PENALTY && !M->pairs.empty())
        return fail(NEWT_ERR_UNSUPPORTED, "penalty contact forces are not built; use the PGS solver");
    ModelT& d = M->dev;
    d.n_links = (int)T.size();
    d.nq = nq;
    d.nv = nv;
    d.n_geoms = (int)M->geoms.size();
    d.n_pairs = (int)M->pairs.size();
    d.cap_c = std::max(cap_c, 1);
    d.cap_r = std::max(4 * cap_c, 1);
    d.n_chain = (int)M->chain.size();
    d.integrator = o.integrator;
    d.mode = o.solver_mode;
    d.iterations = o.iterations;
    d.cone = o.cone;
    d.dt = o.dt;
    d.slop = o.penetration_slop;
    for (int k = 0; k < 3; ++k) d.gravity[k] = o.gravity[k];
    HIP_TRY(hipMalloc(&M->d_links, sizeof(LinkT) * std::max<size_t>(1, M->links.size())));
    HIP_TRY(hipMalloc(&M->d_geoms, sizeof(GeomT) * std::max<size_t>(1, M->hgeoms.size())));
    HIP_TRY(hipMalloc(&M->d_pairs, sizeof(PairT) * std::max<size_t>(1, M->pairs.size())));
    HIP_TRY(hipMalloc(&M->d_chain, sizeof(int) * std::max<size_t>(1, M->chain.size())));
    HIP_TRY(hipMemcpy(M->d_links, M->links.data(), sizeof(LinkT) * M->links.size(), hipMemcpyHostToDevice));
    if (!M->hgeoms.empty())
        HIP_TRY(hipMemcpy(M->d_geoms, M->hgeoms.data(), sizeof(GeomT) * M->hgeoms.size(), hipMemcpyHostToDevice));
    if (!M->pairs.empty())
        HIP_TRY(hipMemcpy(M->d_pairs, M->pairs.data(), sizeof(PairT) * M->pairs.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(M->d_chain, M->chain.data(), sizeof(int) * M->chain.size(), hipMemcpyHostToDevice));
    d.links = M->d_links;
    d.geoms = M->d_geoms;
    d.pairs = M->d_pairs;
    d.chain = M->d_chain;
    return NEWT_OK;
}

Layout make_layout(const ModelT& m) {
    Layout L{};
    int r = 0, dd = 0, ii = 0;
    auto R = [&](int n) { int o = r; r += n; return o; };
    auto D = [&](int n) { int o = dd; dd += n; return o; };
    auto I = [&](int n) { int o = ii; ii += n; return o; };
    int n = m.n_links, nv = std::max(m.nv, 1);
    L.poses = R(7 * n);
    L.lvel = R(6 * n);
    L.xup = R(12 * n);
    L.S = R(6 * n);
    L.v = R(6 * n);
    L.c = R(6 * n);
    L.a = R(6 * n);
    L.ia = R(36 * n);
    L.pa = R(6 * n);
    L.d = R(n);
    L.fs = R(n);
    L.tau = R(nv);
    L.qacc_free = R(nv);
    L.jt = R(3 * nv);
    L.cgeo = R(9 * m.cap_c);
    L.coef = R(2 * nv * m.cap_r);
    L.rreg = R(m.cap_r);
    L.rdiag = R(m.cap_r);
    L.rbias = R(m.cap_r);
    L.f = R(m.cap_r);
    L.A = R(m.cap_r * m.cap_r);
    L.iw = R(n);
    L.delta = R(nv);
    L.gvec = R(nv);
    L.gres = R(nv);
    L.total_r = r;
    L.M = D(nv * nv);
    L.Li = D(nv * nv);
    L.Ld = D(nv * nv);
    L.dvec = D(nv);
    L.dres = D(nv);
    L.total_d = dd;
    L.cids = I(3 * m.cap_c);
    L.rinfo = I(3 * m.cap_r);
    L.pairs = I(std::max(m.n_pairs, 1));
    L.iwf = I(n + 2);
    L.total_i = ii;
    return L;
}

template <class R>
__global__ void __launch_bounds__(kBlock) k_step(ModelT m, Layout L, DataT d, R* ws, double* dws, int* iws,
                                                 int n_steps, int init) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.W) return;
    World<R> wd{m, L, {ws, d.W, w}, {dws, d.W, w}, {iws, d.W, w}, {d.qpos, d.W, w}, {d.qvel, d.W, w},
                {d.qacc, d.W, w}, {d.qfrc, d.W, w}, {d.fat, d.W, w}};
    if (d.fault[w] != 0) return;
    const R dt = (R)m.dt;
    wd.forward_kinematics();
    wd.link_velocities();
    if (init) {
        wd.populate(dt, true);
        iws[(size_t)(L.iwf + m.n_links) * d.W + w] = 0;
        iws[(size_t)(L.iwf + m.n_links + 1) * d.W + w] = 0;
        return;
    }
    // update_broadphase at step start (world.rs:1190-1226): a no-op unless
    // the state was set since the last step.
    wd.populate(dt, false);
    for (int s = 0; s < n_steps; ++s) {
        int st = wd.step();
        if (st != 0) {
            d.fault[w] = st;
            break;
        }
    }
    d.npair[w] = iws[(size_t)(L.iwf + m.n_links) * d.W + w];
    d.ncontact[w] = iws[(size_t)(L.iwf + m.n_links + 1) * d.W + w];
}

newt_status launch(newt_data* D, int n_steps, hipStream_t stream, int init) {
    const newt_model* M = D->model;
    dim3 grid((D->W + kBlock - 1) / kBlock), block(kBlock);
    if (D->precision == NEWT_PRECISION_FP64)
        hipLaunchKernelGGL(k_step<double>, grid, block, 0, stream, M->dev, D->L, D->d, (double*)D->ws, D->dws, D->iws,
                           n_steps, init);
    else
        hipLaunchKernelGGL(k_step<float>, grid, block, 0, stream, M->dev, D->L, D->d, (float*)D->ws, D->dws, D->iws,
                           n_steps, init);
    HIP_TRY(hipGetLastError());
    D->last_stream = stream;
    return NEWT_OK;
}

// Host [W][n] <-> device [n][W].
void to_lanes(const double* src, std::vector<double>& dst, int W, int n) {
    dst.resize((size_t)W * n);
    for (int w = 0; w < W; ++w)
        for (int k = 0; k < n; ++k) dst[(size_t)k * W + w] = src[(size_t)w * n + k];
}
void from_lanes(const std::vector<double>& src, double* dst, int W, int n) {
    for (int w = 0; w < W; ++w)
        for (int k = 0; k < n; ++k) dst[(size_t)w * n + k] = src[(size_t)k * W + w];
}

}  // namespace

extern "C" {

const char* newt_last_error(void) { return g_err.c_str(); }

newt_status newt_model_create(const newt_options* options, newt_model** out) {
    if (!options || !out) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (!(options->dt > 0.0)) return fail(NEWT_ERR_INVALID_CONFIG, "dt must be positive");
    newt_model* m = new newt_model();
    m->opt = *options;
    *out = m;
    return NEWT_OK;
}

newt_status newt_model_add_tree(newt_model* model, const newt_link_desc* links, int32_t n_links, int32_t self_collide,
                                int32_t* tree_index) {
    if (!model || !links || n_links < 1) return fail(NEWT_ERR_INVALID_ARG, "bad tree arguments");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is finalized");
    model->trees.emplace_back(links, links + n_links);
    model->tree_collide.push_back(self_collide);
    if (tree_index) *tree_index = (int32_t)model->trees.size() - 1;
    return NEWT_OK;
}

newt_status newt_model_add_geom(newt_model* model, const newt_geom_desc* geom, int32_t* geom_index) {
    if (!model || !geom) return fail(NEWT_ERR_INVALID_ARG, "bad geom arguments");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is finalized");
    model->geoms.push_back(*geom);
    if (geom_index) *geom_index = (int32_t)model->geoms.size() - 1;
    return NEWT_OK;
}

newt_status newt_model_finalize(newt_model* model) {
    if (!model) return fail(NEWT_ERR_INVALID_ARG, "null model");
    if (model->finalized) return NEWT_OK;
    newt_status s = compile(model);
    if (s != NEWT_OK) return s;
    model->finalized = true;
    return NEWT_OK;
}

newt_status newt_model_dims(const newt_model* model, int32_t* nq, int32_t* nv, int32_t* n_links, int32_t* n_geoms) {
    if (!model || !model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is not finalized");
    if (nq) *nq = model->dev.nq;
    if (nv) *nv = model->dev.nv;
    if (n_links) *n_links = model->dev.n_links;
    if (n_geoms) *n_geoms = model->dev.n_geoms;
    return NEWT_OK;
}

void newt_model_destroy(newt_model* model) {
    if (!model) return;
    (void)hipFree(model->d_links);
    (void)hipFree(model->d_geoms);
    (void)hipFree(model->d_pairs);
    (void)hipFree(model->d_chain);
    delete model;
}

newt_status newt_data_create(const newt_model* model, int32_t n_worlds, int32_t precision, newt_data** out) {
    if (!model || !out || n_worlds < 1) return fail(NEWT_ERR_INVALID_ARG, "bad data arguments");
    if (!model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is not finalized");
    if (precision != NEWT_PRECISION_FP64 && precision != NEWT_PRECISION_MIXED)
        return fail(NEWT_ERR_INVALID_ARG, "unknown precision");
    newt_data* D = new newt_data();
    D->model = model;
    D->W = n_worlds;
    D->precision = precision;
    D->L = make_layout(model->dev);
    const ModelT& m = model->dev;
    size_t W = (size_t)n_worlds;
    size_t rsz = precision == NEWT_PRECISION_FP64 ? sizeof(double) : sizeof(float);
    auto alloc = [&](void** p, size_t bytes) -> bool { return hipMalloc(p, std::max<size_t>(bytes, 8)) == hipSuccess; };
    bool ok = alloc((void**)&D->d.qpos, W * m.nq * 8) && alloc((void**)&D->d.qvel, W * m.nv * 8) &&
              alloc((void**)&D->d.qacc, W * m.nv * 8) && alloc((void**)&D->d.qfrc, W * m.nv * 8) &&
              alloc((void**)&D->d.fat, W * m.n_geoms * 6 * 8) && alloc((void**)&D->d.fault, W * 4) &&
              alloc((void**)&D->d.ncontact, W * 4) && alloc((void**)&D->d.npair, W * 4) &&
              alloc(&D->ws, W * D->L.total_r * rsz) && alloc((void**)&D->dws, W * D->L.total_d * 8) &&
              alloc((void**)&D->iws, W * D->L.total_i * 4);
    if (!ok) {
        newt_data_destroy(D);
        return fail(NEWT_ERR_OOM, "device allocation failed");
    }
    D->d.W = n_worlds;
    std::vector<double> q((size_t)W * m.nq), qd((size_t)W * m.nv, 0.0);
    for (size_t w = 0; w < W; ++w)
        for (int k = 0; k < m.nq; ++k) q[w * m.nq + k] = model->q_default[k];
    HIP_TRY(hipMemset(D->d.qfrc, 0, W * m.nv * 8));
    HIP_TRY(hipMemset(D->d.qacc, 0, W * m.nv * 8));
    HIP_TRY(hipMemset(D->d.fault, 0, W * 4));
    HIP_TRY(hipMemset(D->d.ncontact, 0, W * 4));
    HIP_TRY(hipMemset(D->d.npair, 0, W * 4));
    HIP_TRY(hipMemset(D->iws, 0, W * D->L.total_i * 4));
    *out = D;
    // Default state, then the fat bounds World::add_geom builds from it.
    newt_status s = newt_data_set_state(D, q.data(), qd.data());
    if (s != NEWT_OK) return s;
    s = launch(D, 0, nullptr, 1);
    if (s != NEWT_OK) return s;
    HIP_TRY(hipDeviceSynchronize());
    return NEWT_OK;
}

void newt_data_destroy(newt_data* D) {
    if (!D) return;
    (void)hipFree(D->d.qpos);
    (void)hipFree(D->d.qvel);
    (void)hipFree(D->d.qacc);
    (void)hipFree(D->d.qfrc);
    (void)hipFree(D->d.fat);
    (void)hipFree(D->d.fault);
    (void)hipFree(D->d.ncontact);
    (void)hipFree(D->d.npair);
    (void)hipFree(D->ws);
    (void)hipFree(D->dws);
    (void)hipFree(D->iws);
    delete D;
}

newt_status newt_data_set_state(newt_data* D, const double* qpos, const double* qvel) {
    if (!D || !qpos || !qvel) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    std::vector<double> a, b;
    to_lanes(qpos, a, D->W, m.nq);
    to_lanes(qvel, b, D->W, m.nv);
    HIP_TRY(hipDeviceSynchronize());
    if (m.nq) HIP_TRY(hipMemcpy(D->d.qpos, a.data(), a.size() * 8, hipMemcpyHostToDevice));
    if (m.nv) HIP_TRY(hipMemcpy(D->d.qvel, b.data(), b.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(D->d.fault, 0, (size_t)D->W * 4));
    return NEWT_OK;
}

newt_status newt_data_get_state(const newt_data* D, double* qpos, double* qvel) {
    if (!D) return fail(NEWT_ERR_INVALID_ARG, "null data");
    const ModelT& m = D->model->dev;
    HIP_TRY(hipDeviceSynchronize());
    std::vector<double> a((size_t)D->W * m.nq), b((size_t)D->W * m.nv);
    if (qpos && m.nq) {
        HIP_TRY(hipMemcpy(a.data(), D->d.qpos, a.size() * 8, hipMemcpyDeviceToHost));
        from_lanes(a, qpos, D->W, m.nq);
    }
    if (qvel && m.nv) {
        HIP_TRY(hipMemcpy(b.data(), D->d.qvel, b.size() * 8, hipMemcpyDeviceToHost));
        from_lanes(b, qvel, D->W, m.nv);
    }
    return NEWT_OK;
}

newt_status newt_data_get_qacc(const newt_data* D, double* qacc) {
    if (!D || !qacc) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    HIP_TRY(hipDeviceSynchronize());
    std::vector<double> a((size_t)D->W * m.nv);
    if (m.nv) {
        HIP_TRY(hipMemcpy(a.data(), D->d.qacc, a.size() * 8, hipMemcpyDeviceToHost));
        from_lanes(a, qacc, D->W, m.nv);
    }
    return NEWT_OK;
}

newt_status newt_data_get_faults(const newt_data* D, int32_t* faults) {
    if (!D || !faults) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(faults, D->d.fault, (size_t)D->W * 4, hipMemcpyDeviceToHost));
    return NEWT_OK;
}

newt_status newt_data_contact_capacity(const newt_data* D, int32_t* cap) {
    if (!D || !cap) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    *cap = D->model->dev.cap_c;
    return NEWT_OK;
}

newt_status newt_data_get_contacts(const newt_data* D, int32_t* counts, int32_t* ids, double* geo) {
    if (!D || !counts || !ids || !geo) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    const Layout& L = D->L;
    size_t W = (size_t)D->W;
    int cap = m.cap_c;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(counts, D->d.ncontact, W * 4, hipMemcpyDeviceToHost));
    std::vector<int> ih(W * 3 * cap);
    HIP_TRY(hipMemcpy(ih.data(), D->iws + (size_t)L.cids * W, ih.size() * 4, hipMemcpyDeviceToHost));
    std::vector<double> gh(W * 9 * cap);
    if (D->precision == NEWT_PRECISION_FP64) {
        HIP_TRY(hipMemcpy(gh.data(), (double*)D->ws + (size_t)L.cgeo * W, gh.size() * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> fh(W * 9 * cap);
        HIP_TRY(hipMemcpy(fh.data(), (float*)D->ws + (size_t)L.cgeo * W, fh.size() * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < fh.size(); ++k) gh[k] = fh[k];
    }
    for (size_t w = 0; w < W; ++w)
        for (int k = 0; k < cap; ++k) {
            for (int j = 0; j < 3; ++j) ids[(w * cap + k) * 3 + j] = ih[(size_t)(3 * k + j) * W + w];
            for (int j = 0; j < 7; ++j) geo[(w * cap + k) * 7 + j] = gh[(size_t)(9 * k + j) * W + w];
        }
    return NEWT_OK;
}

newt_status newt_data_pair_capacity(const newt_data* D, int32_t* cap) {
    if (!D || !cap) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    *cap = std::max(D->model->dev.n_pairs, 1);
    return NEWT_OK;
}

newt_status newt_data_get_pairs(const newt_data* D, int32_t* counts, int32_t* pairs) {
    if (!D || !counts || !pairs) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const newt_model* M = D->model;
    size_t W = (size_t)D->W;
    int cap = std::max(M->dev.n_pairs, 1);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(counts, D->d.npair, W * 4, hipMemcpyDeviceToHost));
    std::vector<int> ih(W * cap);
    HIP_TRY(hipMemcpy(ih.data(), D->iws + (size_t)D->L.pairs * W, ih.size() * 4, hipMemcpyDeviceToHost));
    for (size_t w = 0; w < W; ++w)
        for (int k = 0; k < cap; ++k) {
            int idx = k < counts[w] ? ih[(size_t)k * W + w] : 0;
            pairs[(w * cap + k) * 2] = k < counts[w] ? M->pairs[idx].a : -1;
            pairs[(w * cap + k) * 2 + 1] = k < counts[w] ? M->pairs[idx].b : -1;
        }
    return NEWT_OK;
}

newt_status newt_step(const newt_model* model, newt_data* data, int32_t n_steps, void* stream) {
    if (!model || !data || n_steps < 0) return fail(NEWT_ERR_INVALID_ARG, "bad step arguments");
    if (data->model != model) return fail(NEWT_ERR_INVALID_ARG, "data was created for another model");
    if (n_steps == 0) return NEWT_OK;
    return launch(data, n_steps, (hipStream_t)stream, 0);
}

newt_status newt_data_synchronize(newt_data* data) {
    if (!data) return fail(NEWT_ERR_INVALID_ARG, "null data");
    HIP_TRY(hipDeviceSynchronize());
    return NEWT_OK;
}

}  // extern "C"
