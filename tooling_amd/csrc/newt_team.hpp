// newt_team.hpp — persistent wavefront-per-world step.
//
// One 64-lane wavefront steps one world through every step of a launch.  The
// world's state and all intermediates (poses, Plücker transforms, inertias,
// mass-matrix factors, contacts, Jacobian rows) stay in LDS for the whole
// launch; global memory is read once at the start (state, fat bounds) and
// written once at the end (state, qacc, fat bounds, the last step's contacts
// and pairs).
//
// Work inside the wavefront:
//   tree passes (FK, link velocities, CRB, ABA)  one lane per link of a tree
//       level, root to leaves or leaves to root; siblings adding into a common
//       parent take turns by sibling rank, so every sum has a fixed order
//   broadphase refresh                            one lane per geom
//   pair test + narrowphase                       one lane per candidate pair;
//       contacts land in candidate order, then slot order (world.rs:2476-2528)
//   Cholesky                                      one lane per matrix row,
//       right-looking, row in registers, pivots broadcast by v_readlane.  Each
//       entry receives its updates in the same order as the reference's
//       Cholesky-Banachiewicz (dynamics.rs:602-622)
//   diagApprox weights                            one lane per (link, axis)
//   Jacobian rows, bias, regularizer              one lane per row
//   Delassus + PGS                                one lane per row: row i of
//       A = Y Yᵀ (Y = L⁻¹Jᵀ) in registers, residuals updated incrementally
// The contact force reaches the accelerations through the factor of
// M + dt·B:  qacc = qacc_free + (M+dtB)⁻¹ Jᵀ f/dt = qacc_free + L⁻ᵀ (Y f/dt),
// the linear-in-τ identity of the implicit-damping ABA that the reference
// runs a second time (world.rs:1952-2051).
// R is the storage type of the Jacobian rows and of A: double for the FP64
// step, float for the mixed step.  Everything else is FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "newt_math.hpp"
#include "newt_types.hpp"

namespace newt {

// LDS layout of one world.  Doubles live across the whole step (state,
// poses, transforms, factor, contacts); then a union of two scratch sets that
// are never live together: the CRB/ABA set (ia, v, c, a, pa, fs, dd) and the
// contact-solve set (row bias/regularizer/diagonal, impulses, R rows).  The
// ints follow the union.
struct TeamLayout {
    int q, qd, qfrc, qacc, qfree, tau, u, dvec;
    int poses, lvel, xup, Li, Ld, Lr, Ldr, fat, cgeo, invw;
    int ia, v, c, a, pa, fs, dd;  // union, CRB/ABA side
    int rbias, rreg, rdg, f;      // union, solve side (up to 64 rows)
    int lg;                       // union: bias | reg | diag | impulse, cap_r each (over 64 rows)
    int total_d;                  // doubles, union included
    int j_off;                    // byte offset of the rows: 64 x nv of R (union, solve side)
    int i_off;                    // byte offset of the ints
    int cids, ccnt, coff, plist, alist, misc;
    int q0, qd0, aq, av;          // RK4: start state and weighted stage sums
    int qlim, lsg, lvi;           // joint limits: force [nv]; row coefficient, violation
    int ylim;                     // union, solve side: rows of L⁻¹ Eᵀ, FP64, n_ranged x nv
    int lsl, llm;                 // ints: row dof slot, limit index
    int lflag, lused;             // ints: links touched by active contacts (flags, list)
    int bytes;                    // per world, multiple of 16
    // Global block per world holding the rows of contact sets of more than
    // 64 rows (cap_r x nv of R; stride 0 when the model cannot produce them).
    int ov_stride;
};

inline TeamLayout make_team_layout(const ModelT& m, int real_bytes) {
    TeamLayout T{};
    int d = 0;
    auto D = [&](int n) { int o = d; d += n > 0 ? n : 1; return o; };
    const int n = m.n_links, nv = m.nv, nq = m.nq, cr = m.cap_r > 0 ? m.cap_r : 1;
    const int lr = cr < 64 ? cr : 64;  // rows held in LDS
    T.q = D(nq);
    T.qd = D(nv);
    T.qfrc = D(nv);
    T.qacc = D(nv);
    T.qfree = D(nv);
    T.tau = D(nv);
    T.u = D(nv);
    T.dvec = D(nv);
    T.poses = D(7 * n);
    T.lvel = D(6 * n);
    T.xup = D(12 * n);
    T.Li = D(nv * nv);
    T.Ld = m.damped ? D(nv * nv) : T.Li;
    T.Lr = D(nv);
    T.Ldr = m.damped ? D(nv) : T.Lr;
    T.fat = D(6 * m.n_geoms);
    T.cgeo = D(9 * m.cap_c);
    T.invw = D(n);
    const bool rk4 = m.integrator == 2;
    T.q0 = rk4 ? D(nq) : 0;
    T.qd0 = rk4 ? D(nv) : 0;
    T.aq = rk4 ? D(nq) : 0;
    T.av = rk4 ? D(nv) : 0;
    const int nrl = m.n_ranged;
    T.qlim = nrl ? D(nv) : 0;
    T.lsg = nrl ? D(nrl) : 0;
    T.lvi = nrl ? D(nrl) : 0;
    d = (d + 1) / 2 * 2;  // 16-byte aligned union
    const int u0 = d;
    T.ia = D(36 * n);
    T.v = D(6 * n);
    T.c = D(6 * n);
    T.a = D(6 * n);
    T.pa = D(6 * n);
    T.fs = D(n);
    T.dd = D(n);
    const int aba_end = d;
    d = u0;
    T.rbias = D(lr);
    T.rreg = D(lr);
    T.rdg = D(lr);
    T.f = D(64);
    d = (d + 1) / 2 * 2;
    T.j_off = 8 * d;
    const int solve_end_bytes = T.j_off + real_bytes * lr * (nv > 0 ? nv : 1);
    T.lg = u0;
    const int large_end_bytes = cr > 64 ? 8 * (u0 + 4 * cr) : 0;
    T.ylim = u0;
    const int lim_end_bytes = 8 * (u0 + nrl * nv);
    const int union_end_bytes =
        std::max(std::max(std::max(8 * aba_end, solve_end_bytes), large_end_bytes), lim_end_bytes);
    T.total_d = (union_end_bytes + 7) / 8;
    T.i_off = (union_end_bytes + 15) / 16 * 16;
    int i = 0;
    auto I = [&](int k) { int o = i; i += k > 0 ? k : 1; return o; };
    T.cids = I(3 * m.cap_c);
    T.ccnt = I(m.n_pairs);
    T.coff = I(m.n_pairs);
    T.plist = I(m.n_pairs);
    T.alist = I(m.cap_c);
    T.misc = I(8);
    T.lsl = I(nrl);
    T.llm = I(nrl);
    T.lflag = I(n);
    T.lused = I(n);
    T.bytes = (T.i_off + 4 * i + 15) / 16 * 16;
    if (cr > 64) T.ov_stride = (real_bytes * cr * (nv > 0 ? nv : 1) + 255) / 256 * 256;
    return T;
}

// Byte offsets of the model tables as cached in LDS (shared by the worlds of
// a workgroup); the same layout is the device blob they are copied from.
struct CacheLayout {
    int links, geoms, pairs, level_off, level_links, level_nrank, dof_link, limits;
    int bytes;
};

inline CacheLayout make_cache_layout(const ModelT& m) {
    CacheLayout C{};
    int b = 0;
    auto B = [&](int bytes) { int o = b; b += (bytes + 15) / 16 * 16; return o; };
    C.links = B((int)sizeof(LinkT) * m.n_links);
    C.geoms = B((int)sizeof(GeomT) * m.n_geoms);
    C.pairs = B((int)sizeof(PairT) * m.n_pairs);
    C.level_off = B(4 * (m.n_levels + 1));
    C.level_links = B(4 * m.n_links);
    C.level_nrank = B(4 * m.n_levels);
    C.dof_link = B(4 * m.nv);
    C.limits = B((int)sizeof(LimitT) * m.n_ranged);
    C.bytes = b > 0 ? b : 16;
    return C;
}

// Lane `lane`'s value of v in every lane (v_readlane on both halves).
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    int2 p = __builtin_bit_cast(int2, v);
    p.x = __builtin_amdgcn_readlane(p.x, lane);
    p.y = __builtin_amdgcn_readlane(p.y, lane);
    return __builtin_bit_cast(double, p);
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
__device__ __forceinline__ double readlane_r(double v, int lane) { return readlane_f64(v, lane); }
__device__ __forceinline__ float readlane_r(float v, int lane) { return readlane_f32(v, lane); }

// v, with x in lane i.  `lane` passes through an empty asm so the
// compare is evaluated here: otherwise the compiler hoists one lane mask per
// unrolled i out of the enclosing loops and spills them.
template <class P>
__device__ __forceinline__ P put_lane(P v, P x, int lane, int i) {
    asm volatile("" : "+v"(lane));
    return lane == i ? x : v;
}

// v where i < lane, else 0 (same opaque lane as put_lane).
__device__ __forceinline__ double below_lane(double v, int lane, int i) {
    asm volatile("" : "+v"(lane));
    return i < lane ? v : 0.0;
}

// v where i <= lane, else 0 (same opaque lane as put_lane).
__device__ __forceinline__ double upto_lane(double v, int lane, int i) {
    asm volatile("" : "+v"(lane));
    return i <= lane ? v : 0.0;
}

// Orders this wavefront's LDS accesses: a wavefront's LDS operations execute
// in issue order, so a fence for the compiler is all one world needs.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The same for a wavefront's global-memory accesses (waits for them).
__device__ __forceinline__ void gsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ unsigned long long lanes_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Where a step's contact rows live: all in LDS for up to 64 rows; beyond,
// the rows in the world's global overflow block and the per-row scalars in
// LDS.
template <class R> struct RowStore {
    R* J;          // rows x nv
    double* bias;  // per row
    double* reg;
    double* diag;  // Delassus diagonal (+ regularizer, over 64 rows)
    double* f;
};

// The model tables as cached in the workgroup's LDS.
struct Tables {
    const LinkT* links;
    const GeomT* geoms;
    const PairT* pairs;
    const int* level_off;
    const int* level_links;
    const int* level_nrank;
    const int* dof_link;
    const LimitT* limits;
};

// X: the extended step (joint-limit rows, RK4, sphere/capsule and box-box
// pairs).  Models without them run the X = false instantiation, whose
// register allocation is that of the plane-contact Euler step alone.
template <class R, int NV, bool X>
struct Team {
    const ModelT& m;      // scalars of the model (kernel argument)
    const TeamLayout& T;  // kernel argument
    const Tables tb;      // table pointers into LDS
    const int lane;
    double* sd;
    int* si;
    R* sJ;
    RowStore<R> ov;  // this world's overflow block (contact sets over 64 rows)

    __device__ __forceinline__ V3<double> ld3(int o) const { return {sd[o], sd[o + 1], sd[o + 2]}; }
    __device__ __forceinline__ void st3(int o, V3<double> v) const { sd[o] = v.x; sd[o + 1] = v.y; sd[o + 2] = v.z; }
    __device__ __forceinline__ SM<double> ldm(int o) const { return {ld3(o), ld3(o + 3)}; }
    __device__ __forceinline__ void stm(int o, SM<double> v) const { st3(o, v.w); st3(o + 3, v.v); }
    __device__ __forceinline__ SF<double> ldf(int o) const { return {ld3(o), ld3(o + 3)}; }
    __device__ __forceinline__ void stf(int o, SF<double> v) const { st3(o, v.t); st3(o + 3, v.f); }
    __device__ __forceinline__ V3<double> pos(int i) const { return ld3(T.poses + 7 * i); }
    __device__ __forceinline__ Qt<double> ori(int i) const {
        int o = T.poses + 7 * i + 3;
        return {sd[o], sd[o + 1], sd[o + 2], sd[o + 3]};
    }
    __device__ __forceinline__ Xf<double> ldx(int i) const {
        Xf<double> x;
#pragma unroll
        for (int k = 0; k < 9; ++k) x.E.m[k] = sd[T.xup + 12 * i + k];
        x.r = ld3(T.xup + 12 * i + 9);
        return x;
    }
    __device__ __forceinline__ void ld36(int o, double* a) const {
#pragma unroll
        for (int e = 0; e < 36; ++e) a[e] = sd[o + e];
    }
    __device__ __forceinline__ static V3<double> v3of(const double* a) { return mk3(a[0], a[1], a[2]); }

    // ---------------- kinematics (tree.rs:771-858, world.rs:2777-2854) ----
    __device__ __forceinline__ void link_pose(int i) const {
        const LinkT& l = tb.links[i];
        V3<double> p;
        Qt<double> o;
        V3<double> pj = v3of(l.pj), cj = v3of(l.cj);
        if (l.joint == JF_FREE) {
            int qo = T.q + l.q_off;
            p = mk3(sd[qo], sd[qo + 1], sd[qo + 2]);
            o = {sd[qo + 3], sd[qo + 4], sd[qo + 5], sd[qo + 6]};
        } else if (l.joint == JF_FIXED) {
            V3<double> pp = z3<double>();
            Qt<double> po = qid<double>();
            if (l.parent >= 0) {
                pp = pos(l.parent);
                po = ori(l.parent);
            }
            Qt<double> pjq{l.pjq[0], l.pjq[1], l.pjq[2], l.pjq[3]};
            Qt<double> cjq{l.cjq[0], l.cjq[1], l.cjq[2], l.cjq[3]};
            V3<double> jp = pp + qrot(po, pj);
            o = qmul(qmul(po, pjq), qconj(cjq));
            p = jp - qrot(o, cj);
        } else {
            V3<double> pp = pos(l.parent);
            Qt<double> po = ori(l.parent);
            double qq = sd[T.q + l.q_off];
            V3<double> jp = pp + qrot(po, pj);
            if (l.joint == JF_HINGE) {
                o = qmul(po, qaxis(v3of(l.axis), qq));
                p = jp - qrot(o, cj);
            } else {
                o = po;
                p = (jp + qrot(po, v3of(l.axis) * qq)) - qrot(o, cj);
            }
        }
        int b = T.poses + 7 * i;
        sd[b] = p.x;
        sd[b + 1] = p.y;
        sd[b + 2] = p.z;
        sd[b + 3] = o.x;
        sd[b + 4] = o.y;
        sd[b + 5] = o.z;
        sd[b + 6] = o.w;
    }
    // (linear velocity at the COM, angular velocity), world frame.
    __device__ __forceinline__ void link_velocity(int i) const {
        const LinkT& l = tb.links[i];
        V3<double> lin = z3<double>(), ang = z3<double>();
        if (l.parent < 0) {
            if (l.joint == JF_FREE) {
                Qt<double> o = ori(i);
                int vo = T.qd + l.v_off;
                ang = qrot(o, mk3(sd[vo], sd[vo + 1], sd[vo + 2]));
                lin = qrot(o, mk3(sd[vo + 3], sd[vo + 4], sd[vo + 5]));
            }
        } else {
            V3<double> pv = ld3(T.lvel + 6 * l.parent), pw = ld3(T.lvel + 6 * l.parent + 3);
            V3<double> cp = pos(i), pp = pos(l.parent);
            Qt<double> po = ori(l.parent);
            V3<double> pac = pv + cross3(pw, cp - pp);
            if (l.joint == JF_HINGE) {
                V3<double> aw = qrot(po, v3of(l.axis));
                double r = sd[T.qd + l.v_off];
                ang = pw + aw * r;
                V3<double> jw = pp + qrot(po, v3of(l.pj));
                lin = pac + cross3(aw * r, cp - jw);
            } else if (l.joint == JF_SLIDE) {
                lin = pac + qrot(po, v3of(l.axis)) * sd[T.qd + l.v_off];
                ang = pw;
            } else {
                lin = pac;
                ang = pw;
            }
        }
        st3(T.lvel + 6 * i, lin);
        st3(T.lvel + 6 * i + 3, ang);
    }
    __device__ __forceinline__ void kinematics() const {
        for (int lev = 0; lev < m.n_levels; ++lev) {
            for (int k = tb.level_off[lev] + lane; k < tb.level_off[lev + 1]; k += 64) link_pose(tb.level_links[k]);
            wsync();
        }
        for (int lev = 0; lev < m.n_levels; ++lev) {
            for (int k = tb.level_off[lev] + lane; k < tb.level_off[lev + 1]; k += 64) link_velocity(tb.level_links[k]);
            wsync();
        }
    }

    // geom_world_pose (geom.rs:887-898)
    __device__ __forceinline__ void geom_pose(const GeomT& g, V3<double>* p, Qt<double>* o) const {
        V3<double> gp = v3of(g.pos);
        Qt<double> gq{g.quat[0], g.quat[1], g.quat[2], g.quat[3]};
        if (g.link < 0) {
            *p = gp;
            *o = gq;
            return;
        }
        Qt<double> lo = ori(g.link);
        *p = pos(g.link) + qrot(lo, gp);
        *o = qmul(lo, gq);
    }
    // geom_aabb (broadphase.rs:708-770)
    __device__ __forceinline__ static void geom_aabb(const GeomT& g, V3<double> p, Qt<double> o, V3<double>* mn, V3<double>* mx) {
        if (g.shape == SH_PLANE) {
            double inf = __builtin_inf();
            *mn = mk3(-inf, -inf, -inf);
            *mx = mk3(inf, inf, inf);
            return;
        }
        V3<double> e;
        if (g.shape == SH_SPHERE) {
            e = mk3(g.size[0], g.size[0], g.size[0]);
        } else if (g.shape == SH_BOX) {
            e = (abs3(qrot(o, mk3(1.0, 0.0, 0.0))) * g.size[0] + abs3(qrot(o, mk3(0.0, 1.0, 0.0))) * g.size[1]) +
                abs3(qrot(o, mk3(0.0, 0.0, 1.0))) * g.size[2];
        } else {
            e = abs3(qrot(o, mk3(0.0, 0.0, 1.0))) * g.size[1] + mk3(g.size[0], g.size[0], g.size[0]);
        }
        *mn = p - e;
        *mx = p + e;
    }
    // Swept bound + fat-bound hysteresis of each geom (world.rs:1254-1358,
    // broadphase.rs:46-54, 320-346); one lane per geom.
    __device__ __forceinline__ void populate(bool insert_all) const {
        const double dt = m.dt;
        for (int gi = lane; gi < m.n_geoms; gi += 64) {
            const GeomT& g = tb.geoms[gi];
            V3<double> sp, smn, smx, emn, emx;
            Qt<double> so;
            geom_pose(g, &sp, &so);
            geom_aabb(g, sp, so, &smn, &smx);
            double ang_speed = 0.0;
            V3<double> pivot = sp;
            if (g.link >= 0) {
                V3<double> v = ld3(T.lvel + 6 * g.link), w = ld3(T.lvel + 6 * g.link + 3);
                V3<double> lp = pos(g.link);
                Qt<double> lo = ori(g.link);
                V3<double> ep = lp + v * dt;
                Qt<double> eo = qintegrate(lo, qrot_inv(lo, w), dt);
                Qt<double> gq{g.quat[0], g.quat[1], g.quat[2], g.quat[3]};
                geom_aabb(g, ep + qrot(eo, v3of(g.pos)), qmul(eo, gq), &emn, &emx);
                ang_speed = len3(w);
                pivot = lp;
            } else {
                emn = smn;
                emx = smx;
            }
            V3<double> bc = (smn + smx) * 0.5;
            double br = len3((smx - smn) * 0.5);
            double pr = len3(bc - pivot) + br;
            double sweep = (ang_speed > 0.0 && nfinite(pr)) ? pr * ang_speed * dt : 0.0;
            V3<double> mn = mk3(nmin(smn.x, emn.x), nmin(smn.y, emn.y), nmin(smn.z, emn.z));
            V3<double> mx = mk3(nmax(smx.x, emx.x), nmax(smx.y, emx.y), nmax(smx.z, emx.z));
            mn = (mn - mk3(sweep, sweep, sweep)) - mk3(g.margin, g.margin, g.margin);
            mx = (mx + mk3(sweep, sweep, sweep)) + mk3(g.margin, g.margin, g.margin);
            double* f = sd + T.fat + 6 * gi;
            bool inside = f[0] <= mn.x && f[1] <= mn.y && f[2] <= mn.z && f[3] >= mx.x && f[4] >= mx.y && f[5] >= mx.z;
            if (!inside || insert_all) {
                V3<double> e = mx - mn;
                V3<double> fm = mk3(nmax(e.x * 0.1, 1.0e-4), nmax(e.y * 0.1, 1.0e-4), nmax(e.z * 0.1, 1.0e-4));
                V3<double> a = mn - fm, b = mx + fm;
                f[0] = a.x;
                f[1] = a.y;
                f[2] = a.z;
                f[3] = b.x;
                f[4] = b.y;
                f[5] = b.z;
            }
        }
        wsync();
    }

    // ---------------- pairs + narrowphase (world.rs:1173-1226, contact.rs:269-403)
    // Contacts of candidate pair k, -1 when its fat bounds do not overlap.
    // With out >= 0 they are written from contact slot `out` on.
    __device__ __forceinline__ int pair_contacts(int k, int out) const {
        const PairT& pr = tb.pairs[k];
        const double* fa = sd + T.fat + 6 * pr.a;
        const double* fb = sd + T.fat + 6 * pr.b;
        bool ov = fa[0] <= fb[3] && fa[3] >= fb[0] && fa[1] <= fb[4] && fa[4] >= fb[1] && fa[2] <= fb[5] &&
                  fa[5] >= fb[2];
        if (!ov) return -1;
        const GeomT& ga = tb.geoms[pr.a];
        const GeomT& gb = tb.geoms[pr.b];
        double fr = ga.friction < gb.friction ? ga.friction : gb.friction;
        double margin = ga.margin > gb.margin ? ga.margin : gb.margin;
        double gap = ga.gap > gb.gap ? ga.gap : gb.gap;
        if (m.slop > 0.0 && gap < m.slop) gap = m.slop;  // append_contacts (world.rs:2517-2528)
        if constexpr (X) {
        if (ga.shape == SH_BOX && gb.shape == SH_BOX) {
            V3<double> pa, pb, nrm, cp[4];
            Qt<double> qa, qb;
            double pen[4];
            geom_pose(ga, &pa, &qa);
            geom_pose(gb, &pb, &qb);
            const int c = box_box(pa, qa, ga.size, pb, qb, gb.size, margin, cp, pen, &nrm);
            if (out >= 0)
                for (int i = 0; i < c; ++i) {
                    const int o = out + i;
                    si[T.cids + 3 * o] = pr.a;
                    si[T.cids + 3 * o + 1] = pr.b;
                    si[T.cids + 3 * o + 2] = i;
                    double* g = sd + T.cgeo + 9 * o;
                    g[0] = cp[i].x;
                    g[1] = cp[i].y;
                    g[2] = cp[i].z;
                    g[3] = nrm.x;
                    g[4] = nrm.y;
                    g[5] = nrm.z;
                    g[6] = pen[i];
                    g[7] = fr;
                    g[8] = gap;
                }
            return c;
        }
        if (ga.shape != SH_PLANE && gb.shape != SH_PLANE) {
            V3<double> pa, pb, pt, nrm;
            Qt<double> qa, qb;
            double pen;
            geom_pose(ga, &pa, &qa);
            geom_pose(gb, &pb, &qb);
            if (!round_pair(ga.shape == SH_CAPSULE, pa, qa, ga.size, gb.shape == SH_CAPSULE, pb, qb, gb.size, margin,
                            &pt, &nrm, &pen))
                return 0;
            if (out >= 0) {
                si[T.cids + 3 * out] = pr.a;
                si[T.cids + 3 * out + 1] = pr.b;
                si[T.cids + 3 * out + 2] = 0;
                double* g = sd + T.cgeo + 9 * out;
                g[0] = pt.x;
                g[1] = pt.y;
                g[2] = pt.z;
                g[3] = nrm.x;
                g[4] = nrm.y;
                g[5] = nrm.z;
                g[6] = pen;
                g[7] = fr;
                g[8] = gap;
            }
            return 1;
        }
        }
        // (primitive, plane) as given, else swapped with the normal negated
        // (contact.rs:2579-2588).
        bool flip = gb.shape != SH_PLANE;
        const GeomT& gp = flip ? gb : ga;
        const GeomT& gq = flip ? ga : gb;
        V3<double> pp, qp;
        Qt<double> po, qo;
        geom_pose(gq, &pp, &po);
        geom_pose(gp, &qp, &qo);
        V3<double> n = qrot(po, mk3(0.0, 0.0, 1.0));
        V3<double> nn = flip ? -n : n;
        int cnt = 0;
        auto emit = [&](V3<double> p, double pen) {
            if (out >= 0) {
                int c = out + cnt;
                si[T.cids + 3 * c] = pr.a;
                si[T.cids + 3 * c + 1] = pr.b;
                si[T.cids + 3 * c + 2] = cnt;
                double* g = sd + T.cgeo + 9 * c;
                g[0] = p.x;
                g[1] = p.y;
                g[2] = p.z;
                g[3] = nn.x;
                g[4] = nn.y;
                g[5] = nn.z;
                g[6] = pen;
                g[7] = fr;
                g[8] = gap;
            }
            ++cnt;
        };
        if (gp.shape == SH_SPHERE) {
            double raw = dot3(qp - pp, n) - gp.size[0];
            if (raw <= margin) emit(qp - n * (gp.size[0] + raw * 0.5), margin - raw);
        } else if (gp.shape == SH_BOX) {
            double dist = dot3(qp - pp, n);
            for (int i = 0; i < 8; ++i) {
                V3<double> loc = mk3((i & 1) == 0 ? -gp.size[0] : gp.size[0], (i & 2) == 0 ? -gp.size[1] : gp.size[1],
                                     (i & 4) == 0 ? -gp.size[2] : gp.size[2]);
                V3<double> co = qrot(qo, loc);
                double ld = dot3(n, co);
                double raw = dist + ld;
                if (raw > margin || ld > 0.0 || cnt >= 4) continue;
                emit((qp + co) - n * (raw * 0.5), margin - raw);
            }
        } else if (gp.shape == SH_CAPSULE) {
            V3<double> aw = qrot(qo, mk3(0.0, 0.0, 1.0));
            for (int e = 0; e < 2; ++e) {
                V3<double> end = e == 0 ? qp + aw * gp.size[1] : qp - aw * gp.size[1];
                double raw = dot3(end - pp, n) - gp.size[0];
                if (raw <= margin) emit(end - n * (gp.size[0] + raw * 0.5), margin - raw);
            }
        }
        return cnt;
    }
    // Contact count of this step; si[misc] = active pair count, si[plist..]
    // the active candidate indices in order.
    __device__ __forceinline__ int contacts() const {
        for (int k = lane; k < m.n_pairs; k += 64) si[T.ccnt + k] = pair_contacts(k, -1);
        wsync();
        if (lane == 0) {
            int off = 0, np = 0;
            for (int k = 0; k < m.n_pairs; ++k) {
                int c = si[T.ccnt + k];
                si[T.coff + k] = off;
                if (c >= 0) {
                    si[T.plist + np++] = k;
                    off += c;
                }
            }
            si[T.misc] = np;
            si[T.misc + 1] = off;
        }
        wsync();
        // The compiled capacity bounds every pair's count; a larger total is
        // reported as a fault by step() and nothing past it is written.
        if (__builtin_amdgcn_readfirstlane(si[T.misc + 1]) > m.cap_c) return m.cap_c + 1;
        for (int k = lane; k < m.n_pairs; k += 64)
            if (si[T.ccnt + k] > 0) pair_contacts(k, si[T.coff + k]);
        wsync();
        return __builtin_amdgcn_readfirstlane(si[T.misc + 1]);
    }
    // Contacts with penetration beyond the gap (solver.rs:2805-2830) into
    // si[alist..], in order; returns their count.
    __device__ __forceinline__ int active_contacts(int nc) const {
        int na = 0;
        for (int base = 0; base < nc; base += 64) {
            int k = base + lane;
            bool act = k < nc && (sd[T.cgeo + 9 * k + 6] - sd[T.cgeo + 9 * k + 8] > 0.0);
            unsigned long long mask = __ballot(act);
            if (act) si[T.alist + na + __popcll(mask & lanes_below(lane))] = k;
            na += __popcll(mask);
        }
        wsync();
        return na;
    }

    // Links that carry a geom of an active contact, in link order, into
    // si[T.lused..]; returns their count.  The rows read the inverse weights
    // of these links only.
    __device__ __forceinline__ int contact_links(int na) const {
        for (int i = lane; i < m.n_links; i += 64) si[T.lflag + i] = 0;
        wsync();
        for (int k = lane; k < na; k += 64) {
            const int c = si[T.alist + k];
            const int la = tb.geoms[si[T.cids + 3 * c]].link, lb = tb.geoms[si[T.cids + 3 * c + 1]].link;
            if (la >= 0) si[T.lflag + la] = 1;
            if (lb >= 0) si[T.lflag + lb] = 1;
        }
        wsync();
        int nu = 0;
        for (int i0 = 0; i0 < m.n_links; i0 += 64) {
            const int i = i0 + lane;
            const bool used = i < m.n_links && si[T.lflag + i] != 0;
            const unsigned long long bal = __ballot(used);
            if (used) si[T.lused + nu + __popcll(bal & lanes_below(lane))] = i;
            nu += __popcll(bal);
        }
        wsync();
        return __builtin_amdgcn_readfirstlane(nu);
    }

    // ---------------- dynamics ---------------------------------------------
    // Parent-to-child Plücker transform (aba.rs:762-791).
    __device__ __forceinline__ Xf<double> joint_xup(int i) const {
        const LinkT& l = tb.links[i];
        Xf<double> x;
        if (l.joint == JF_FIXED) {
#pragma unroll
            for (int k = 0; k < 9; ++k) x.E.m[k] = l.xfE[k];
            x.r = v3of(l.xfr);
        } else if (l.joint == JF_HINGE) {
            V3<double> ax = unit3(v3of(l.axis));
            double s, c;
            nsincos((0.0 - sd[T.q + l.q_off]) * 0.5, &s, &c);
            x.E = qmat(Qt<double>{ax.x * s, ax.y * s, ax.z * s, c});
            x.r = v3of(l.cj) - mv(x.E, v3of(l.pj));
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) x.E.m[k] = (k % 4 == 0) ? 1.0 : 0.0;
            x.r = l.joint == JF_SLIDE ? (v3of(l.cj) - v3of(l.pj)) - v3of(l.axis) * sd[T.q + l.q_off] : z3<double>();
        }
        return x;
    }
    __device__ __forceinline__ SM<double> subspace(int i) const {
        const LinkT& l = tb.links[i];
        V3<double> ax = v3of(l.axis);
        if (l.joint == JF_HINGE) return {ax, cross3(v3of(l.cj), ax)};
        return {z3<double>(), ax};
    }
    __device__ __forceinline__ void store_xups() const {
        for (int i = lane; i < m.n_links; i += 64) {
            Xf<double> x = joint_xup(i);
#pragma unroll
            for (int k = 0; k < 9; ++k) sd[T.xup + 12 * i + k] = x.E.m[k];
            st3(T.xup + 12 * i + 9, x.r);
        }
        wsync();
    }

    // Column c of Xᵀ A X (m6pullback, spatial.rs:457-481) added into the
    // row-major 6x6 at sd[dst]: the unit motion e_c through X, times A,
    // back through Xᵀ as a force.  Six lanes, one per column, pull a link.
    __device__ __forceinline__ void pull_col_add(const double* a, const Xf<double>& x, int c, int dst) const {
        SM<double> e{mk3(c == 0 ? 1.0 : 0.0, c == 1 ? 1.0 : 0.0, c == 2 ? 1.0 : 0.0),
                     mk3(c == 3 ? 1.0 : 0.0, c == 4 ? 1.0 : 0.0, c == 5 ? 1.0 : 0.0)};
        const SF<double> f = xtforce(x, m6mul(a, xmotion(x, e)));
        sd[dst + 0 * 6 + c] += f.t.x;
        sd[dst + 1 * 6 + c] += f.t.y;
        sd[dst + 2 * 6 + c] += f.t.z;
        sd[dst + 3 * 6 + c] += f.f.x;
        sd[dst + 4 * 6 + c] += f.f.y;
        sd[dst + 5 * 6 + c] += f.f.z;
    }

    // dynamics::mass_matrix (dynamics.rs:474-578), composite rigid bodies,
    // into sd[Li] (row-major nv x nv).
    __device__ __forceinline__ void mass_matrix() const {
        const int nv = m.nv, n = m.n_links;
        for (int k = lane; k < nv * nv; k += 64) sd[T.Li + k] = 0.0;
        for (int k = lane; k < n * 36; k += 64) sd[T.ia + k] = tb.links[k / 36].si[k % 36];
        wsync();
        for (int lev = m.n_levels - 1; lev >= 1; --lev) {
            const int k0 = tb.level_off[lev], cnt = tb.level_off[lev + 1] - k0;
            for (int rank = 0; rank < tb.level_nrank[lev]; ++rank) {
                for (int idx = lane; idx < 6 * cnt; idx += 64) {
                    const int i = tb.level_links[k0 + idx / 6];
                    const LinkT& l = tb.links[i];
                    if (l.rank != rank) continue;
                    double a[36];
                    ld36(T.ia + 36 * i, a);
                    pull_col_add(a, ldx(i), idx % 6, T.ia + 36 * l.parent);
                }
                wsync();
            }
        }
        if (tb.links[0].joint == JF_FREE) {
            int o = tb.links[0].v_off;
            for (int e = lane; e < 36; e += 64) sd[T.Li + (o + e / 6) * nv + (o + e % 6)] = sd[T.ia + e];
        }
        for (int i = 1 + lane; i < n; i += 64) {
            const LinkT& l = tb.links[i];
            if (l.joint == JF_FIXED) continue;
            int vi = l.v_off;
            double a[36];
            ld36(T.ia + 36 * i, a);
            SF<double> f = m6mul(a, subspace(i));
            sd[T.Li + vi * nv + vi] = sdot(subspace(i), f) + l.armature;
            int jc = i;
            while (tb.links[jc].parent >= 0) {
                int j = tb.links[jc].parent;
                f = xtforce(ldx(jc), f);
                jc = j;
                const LinkT& pl = tb.links[j];
                int vj = pl.v_off;
                if (pl.joint == JF_HINGE || pl.joint == JF_SLIDE) {
                    double val = sdot(subspace(j), f);
                    sd[T.Li + vj * nv + vi] = val;
                    sd[T.Li + vi * nv + vj] = val;
                } else if (pl.joint == JF_FREE) {
                    double fc[6] = {f.t.x, f.t.y, f.t.z, f.f.x, f.f.y, f.f.z};
                    for (int kk = 0; kk < 6; ++kk) {
                        sd[T.Li + (vj + kk) * nv + vi] = fc[kk];
                        sd[T.Li + vi * nv + (vj + kk)] = fc[kk];
                    }
                }
            }
        }
        wsync();
    }

    // Cholesky of the symmetric matrix in sd[Li] into its lower factor at
    // sd[dst] (may be Li itself) and the pivots' reciprocals at sd[rdst];
    // lane k also keeps row k in a[] and 1/L_kk in rinv.
    // Right-looking, lane i holding row i in registers: at column j the pivot
    // comes from lane j by v_readlane and every lane updates its row; an
    // entry receives its updates in the same order as the reference's
    // Cholesky-Banachiewicz (dynamics.rs:602-622).  Lanes also carry values
    // above their diagonal; the stored factor has them zeroed.  Returns
    // false, uniformly, when a pivot is not positive.
    __device__ __forceinline__ bool factor(int dst, int rdst, double (&a)[NV], double& rinv) const {
        const int n = m.nv;
        const int row = lane < n ? lane : n - 1;
        rinv = 0.0;
#pragma unroll
        for (int s = 0; s < NV; ++s) a[s] = s < n ? sd[T.Li + row * n + s] : 0.0;
        bool ok = true;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            if (j < n) {
                const double djj = readlane_f64(a[j], j);
                ok = ok && djj > 0.0;
                const double piv = __builtin_sqrt(djj);
                const double rp = 1.0 / piv;
                a[j] = put_lane(a[j] * rp, piv, lane, j);
                rinv = put_lane(rinv, rp, lane, j);
#pragma unroll
                for (int k = j + 1; k < NV; ++k)
                    if (k < n) a[k] = a[k] - a[j] * readlane_f64(a[j], k);
            }
        }
        // Stored with the strict upper triangle zeroed (substitutions read
        // whole rows).
        if (lane < n) {
#pragma unroll
            for (int s = 0; s < NV; ++s)
                if (s < n) sd[dst + lane * n + s] = upto_lane(a[s], lane, s);
            sd[rdst + lane] = rinv;
        }
        wsync();
        return ok;
    }

    // Adds dt·damping to the diagonal of sd[Li] (implicit_mass_matrix,
    // tree.rs:645-681).
    __device__ __forceinline__ void add_damping() const {
        for (int s = lane; s < m.nv; s += 64) sd[T.Li + s * m.nv + s] = sd[T.Li + s * m.nv + s] + m.dt * sd[T.dvec + s];
        wsync();
    }

    // ABA (aba.rs:803-1102) with generalized force sd[tau] and implicit
    // joint damping (Euler); the free root reads tau under the constraint
    // solver (tree.rs:1021-1116).  qddot into sd[out].  Needs store_xups().
    __device__ __forceinline__ void aba(int out, bool implicit) const {
        const double dt = m.dt;
        V3<double> g = v3of(m.gravity);
        // velocities, bias forces, rigid inertias: root -> leaves
        for (int lev = 0; lev < m.n_levels; ++lev) {
            for (int k = tb.level_off[lev] + lane; k < tb.level_off[lev + 1]; k += 64) {
                int i = tb.level_links[k];
                const LinkT& l = tb.links[i];
                SM<double> v, c = smz<double>();
                if (l.joint == JF_FREE) {
                    int o = T.qd + l.v_off;
                    v = {mk3(sd[o], sd[o + 1], sd[o + 2]), mk3(sd[o + 3], sd[o + 4], sd[o + 5])};
                } else if (l.joint == JF_FIXED) {
                    v = xmotion(ldx(i), l.parent >= 0 ? ldm(T.v + 6 * l.parent) : smz<double>());
                } else {
                    SM<double> sq = subspace(i) * sd[T.qd + l.v_off];
                    v = xmotion(ldx(i), ldm(T.v + 6 * l.parent)) + sq;
                    c = crm(v, sq);
                }
                stm(T.v + 6 * i, v);
                stm(T.c + 6 * i, c);
                V3<double> fb = qrot_inv(ori(i), g * l.mass);
                double a[36];
#pragma unroll
                for (int e = 0; e < 36; ++e) {
                    a[e] = l.si[e];
                    sd[T.ia + 36 * i + e] = a[e];
                }
                stf(T.pa + 6 * i, crf(v, m6mul(a, v)) - SF<double>{z3<double>(), fb});
            }
            wsync();
        }
        // articulated inertias: leaves -> root, siblings by rank; six lanes
        // per link, one per column of the pulled-back inertia
        for (int lev = m.n_levels - 1; lev >= 1; --lev) {
            const int k0 = tb.level_off[lev], cnt = tb.level_off[lev + 1] - k0;
            for (int rank = 0; rank < tb.level_nrank[lev]; ++rank) {
                for (int idx = lane; idx < 6 * cnt; idx += 64) {
                    const int i = tb.level_links[k0 + idx / 6], col = idx % 6;
                    const LinkT& l = tb.links[i];
                    if (l.rank != rank) continue;
                    double a[36];
                    ld36(T.ia + 36 * i, a);
                    const Xf<double> x = ldx(i);
                    const SF<double> pa = ldf(T.pa + 6 * i);
                    const SM<double> c = ldm(T.c + 6 * i);
                    SF<double> paf;
                    if (l.joint == JF_FIXED) {
                        paf = pa + m6mul(a, c);
                    } else {
                        const SM<double> s = subspace(i);
                        const SF<double> ias = m6mul(a, s);
                        const double d = (sdot(s, ias) + l.armature) + (implicit ? dt * (l.damping + 0.0) : 0.0);
                        const SF<double> ps = pa + m6mul(a, c);
                        const double fs = (sd[T.tau + l.v_off] + 0.0) - l.damping * sd[T.qd + l.v_off];
                        const double qdd = (fs - sdot(s, ps)) / d;
                        paf = ps + ias * qdd;
                        const double inv_d = 1.0 / d;
                        const double iv[6] = {ias.t.x, ias.t.y, ias.t.z, ias.f.x, ias.f.y, ias.f.z};
#pragma unroll
                        for (int r = 0; r < 6; ++r)
#pragma unroll
                            for (int cc = 0; cc < 6; ++cc) a[r * 6 + cc] -= (iv[r] * iv[cc]) * inv_d;
                        if (col == 0) {
                            sd[T.fs + i] = fs;
                            sd[T.dd + i] = d;
                        }
                    }
                    pull_col_add(a, x, col, T.ia + 36 * l.parent);
                    if (col == 0) stf(T.pa + 6 * l.parent, ldf(T.pa + 6 * l.parent) + xtforce(x, paf));
                }
                wsync();
            }
        }
        // root: free base solves its 6x6 system (GMat6::solve, aba.rs:614-665)
        for (int k = lane; k < m.nv; k += 64) sd[out + k] = 0.0;
        if (lane == 0) {
            SM<double> a0 = smz<double>();
            const LinkT& l = tb.links[0];
            if (l.joint == JF_FREE) {
                const int o = l.v_off;
                double fr[6];
                for (int k = 0; k < 6; ++k) fr[k] = m.mode == 1 ? 0.0 + sd[T.tau + o + k] : 0.0;
                SF<double> rhs = SF<double>{mk3(fr[0], fr[1], fr[2]), mk3(fr[3], fr[4], fr[5])} - ldf(T.pa);
                double A[6][7];
                for (int r = 0; r < 6; ++r)
                    for (int cc = 0; cc < 6; ++cc) A[r][cc] = sd[T.ia + r * 6 + cc];
                const double fd = l.free_damping;
                if (fd != 0.0) {
                    const int q0 = T.qd + o;
                    SM<double> qr{mk3(sd[q0], sd[q0 + 1], sd[q0 + 2]), mk3(sd[q0 + 3], sd[q0 + 4], sd[q0 + 5])};
                    rhs = rhs - SF<double>{qr.w * fd, qr.v * fd};
                    for (int k = 0; k < 6; ++k) A[k][k] += implicit ? dt * fd : 0.0;
                }
                double rv[6] = {rhs.t.x, rhs.t.y, rhs.t.z, rhs.f.x, rhs.f.y, rhs.f.z};
                for (int r = 0; r < 6; ++r) A[r][6] = rv[r];
                // Pivot = last row of largest |a| in the column; the swap is
                // predicated so every index stays a compile-time constant.
#pragma unroll
                for (int col = 0; col < 6; ++col) {
                    int piv = col;
                    double best = nabs(A[col][col]);
#pragma unroll
                    for (int r = col + 1; r < 6; ++r) {
                        const double v = nabs(A[r][col]);
                        if (!(best > v)) {
                            piv = r;
                            best = v;
                        }
                    }
#pragma unroll
                    for (int r = col + 1; r < 6; ++r) {
                        const bool sw = r == piv;
#pragma unroll
                        for (int k = 0; k < 7; ++k) {
                            const double t = A[col][k];
                            A[col][k] = sw ? A[r][k] : t;
                            A[r][k] = sw ? t : A[r][k];
                        }
                    }
                    const double inv = 1.0 / A[col][col];
#pragma unroll
                    for (int r = 0; r < 6; ++r) {
                        if (r == col) continue;
                        const double factor = A[r][col] * inv;
#pragma unroll
                        for (int k = col; k < 7; ++k) A[r][k] -= factor * A[col][k];
                    }
                }
                double x6[6];
                for (int r = 0; r < 6; ++r) x6[r] = A[r][6] / A[r][r];
                a0 = {mk3(x6[0], x6[1], x6[2]), mk3(x6[3], x6[4], x6[5])};
                for (int k = 0; k < 6; ++k) sd[out + o + k] = x6[k];
            }
            stm(T.a, a0);
        }
        wsync();
        // accelerations: root -> leaves
        for (int lev = 1; lev < m.n_levels; ++lev) {
            for (int k = tb.level_off[lev] + lane; k < tb.level_off[lev + 1]; k += 64) {
                int i = tb.level_links[k];
                const LinkT& l = tb.links[i];
                SM<double> ap = xmotion(ldx(i), ldm(T.a + 6 * l.parent));
                if (l.joint == JF_FIXED) {
                    stm(T.a + 6 * i, ap);
                    continue;
                }
                SM<double> prime = ap + ldm(T.c + 6 * i);
                double a[36];
                ld36(T.ia + 36 * i, a);
                SF<double> inner = m6mul(a, prime) + ldf(T.pa + 6 * i);
                SM<double> s = subspace(i);
                double qdd = (sd[T.fs + i] - sdot(s, inner)) / sd[T.dd + i];
                stm(T.a + 6 * i, prime + s * qdd);
                sd[out + l.v_off] = qdd;
            }
            wsync();
        }
    }

    // Right-hand side of the free acceleration, rhs = τ − D q̇ − C(q, q̇), into
    // sd[out]: C by recursive Newton-Euler with q̈ = 0 and gravity as each
    // link's body force (the quantities of the ABA's first pass), forces
    // summed leaves -> root by sibling rank.  With M + dt·D = L Lᵀ,
    // (M + dt·D) q̈_free = rhs is the implicit-damping ABA's system
    // (aba.rs:803-1102, tree.rs:1021-1116).  Needs store_xups().
    __device__ __forceinline__ void rne_rhs(int out) const {
        V3<double> g = v3of(m.gravity);
        for (int lev = 0; lev < m.n_levels; ++lev) {
            for (int k = tb.level_off[lev] + lane; k < tb.level_off[lev + 1]; k += 64) {
                const int i = tb.level_links[k];
                const LinkT& l = tb.links[i];
                SM<double> v, acc = smz<double>();
                if (l.joint == JF_FREE) {
                    const int o = T.qd + l.v_off;
                    v = {mk3(sd[o], sd[o + 1], sd[o + 2]), mk3(sd[o + 3], sd[o + 4], sd[o + 5])};
                } else {
                    const Xf<double> x = ldx(i);
                    const SM<double> vp = l.parent >= 0 ? ldm(T.v + 6 * l.parent) : smz<double>();
                    const SM<double> ap = l.parent >= 0 ? ldm(T.a + 6 * l.parent) : smz<double>();
                    v = xmotion(x, vp);
                    acc = xmotion(x, ap);
                    if (l.joint != JF_FIXED) {
                        const SM<double> sq = subspace(i) * sd[T.qd + l.v_off];
                        v = v + sq;
                        acc = acc + crm(v, sq);
                    }
                }
                stm(T.v + 6 * i, v);
                stm(T.a + 6 * i, acc);
                const V3<double> fb = qrot_inv(ori(i), g * l.mass);
                stf(T.pa + 6 * i, (m6mul(l.si, acc) + crf(v, m6mul(l.si, v))) - SF<double>{z3<double>(), fb});
            }
            wsync();
        }
        for (int lev = m.n_levels - 1; lev >= 1; --lev) {
            for (int rank = 0; rank < tb.level_nrank[lev]; ++rank) {
                for (int k = tb.level_off[lev] + lane; k < tb.level_off[lev + 1]; k += 64) {
                    const int i = tb.level_links[k];
                    const LinkT& l = tb.links[i];
                    if (l.rank != rank) continue;
                    stf(T.pa + 6 * l.parent, ldf(T.pa + 6 * l.parent) + xtforce(ldx(i), ldf(T.pa + 6 * i)));
                }
                wsync();
            }
        }
        for (int i = lane; i < m.n_links; i += 64) {
            const LinkT& l = tb.links[i];
            const SF<double> f = ldf(T.pa + 6 * i);
            if (l.joint == JF_FREE) {
                const int o = l.v_off;
                const double fc[6] = {f.t.x, f.t.y, f.t.z, f.f.x, f.f.y, f.f.z};
                for (int k = 0; k < 6; ++k) {
                    const double tau = m.mode == 1 ? 0.0 + sd[T.tau + o + k] : 0.0;
                    sd[out + o + k] = (tau - l.free_damping * sd[T.qd + o + k]) - fc[k];
                }
            } else if (l.joint != JF_FIXED) {
                sd[out + l.v_off] = ((sd[T.tau + l.v_off] + 0.0) - l.damping * sd[T.qd + l.v_off]) - sdot(subspace(i), f);
            }
        }
        wsync();
    }

    // L⁻¹ x for one vector, x_k in lane k; L by rows in a[] (below the
    // diagonal; the registers above it are not zeroed), 1/L_kk in rinv.
    __device__ __forceinline__ double fwd_vec(double x, const double (&a)[NV], double rinv) const {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if (k < m.nv) {
                const double xk = readlane_f64(x, k) * readlane_f64(rinv, k);
                x = put_lane(x - below_lane(a[k], lane, k) * xk, xk, lane, k);
            }
        }
        return x;
    }

    // L⁻ᵀ x for one vector, x_k in lane k; L from sd[T.Li] (upper triangle
    // zero), 1/L_kk from sd[T.Lr].
    __device__ __forceinline__ double bwd_vec(double x) const {
        const int nv = m.nv;
#pragma unroll
        for (int k = NV - 1; k >= 0; --k) {
            if (k < nv) {
                const double zk = readlane_f64(x, k) * sd[T.Lr + k];
                // L[k][lane] is zero for lane > k: finished lanes keep x.
                x = put_lane(x - sd[T.Li + k * nv + (lane < nv ? lane : 0)] * zk, zk, lane, k);
            }
        }
        return x;
    }

    // Translational Jacobian entry of dof s for a world point, along world
    // direction `dir` (jacobian.rs:44-110); zero when dof s is not on the
    // chain of the link whose ancestor mask is (anc0, anc1).
    __device__ __forceinline__ double jac_entry(int s, unsigned long long anc0, unsigned long long anc1, V3<double> point,
                                                V3<double> dir) const {
        const int j = tb.dof_link[s];
        if (!(((j < 64 ? anc0 : anc1) >> (j & 63)) & 1ull)) return 0.0;
        const LinkT& l = tb.links[j];
        if (l.joint == JF_FREE) {
            const int k = s - l.v_off;
            const int b = k < 3 ? k : k - 3;
            V3<double> e = qrot(ori(j), mk3(b == 0 ? 1.0 : 0.0, b == 1 ? 1.0 : 0.0, b == 2 ? 1.0 : 0.0));
            return k < 3 ? dot3(cross3(e, point - pos(j)), dir) : dot3(e, dir);
        }
        Qt<double> po = ori(l.parent);
        V3<double> aw = qrot(po, v3of(l.axis));
        if (l.joint == JF_HINGE) return dot3(cross3(aw, point - (pos(l.parent) + qrot(po, v3of(l.pj)))), dir);
        return dot3(aw, dir);
    }

    // tree_body_invweight0 (solver.rs:3719-3754), translational part, for
    // every link: trace(J M⁻¹ Jᵀ)/3 at the link origin.  With M = L Lᵀ the
    // trace is Σ_axis ‖L⁻¹ J_axisᵀ‖², so one forward substitution per axis
    // suffices.  One lane per (link, axis); L in sd[ld], 1/L_kk in sd[ldr].
    __device__ __forceinline__ void invweights(int ld, int ldr, int nu) const {
        const int n = nu, nv = m.nv;
        // Passes of 63 lanes keep each link's three axes in one pass.
        for (int base = 0; base < 3 * n; base += 63) {
            const int idx = base + lane;
            const bool on = lane < 63 && idx < 3 * n;
            const int i = on ? si[T.lused + idx / 3] : 0, ax = idx % 3;
            const unsigned long long anc0 = tb.links[i].anc[0], anc1 = tb.links[i].anc[1];
            const V3<double> point = pos(i) + qrot(ori(i), z3<double>());
            const V3<double> dir = mk3(ax == 0 ? 1.0 : 0.0, ax == 1 ? 1.0 : 0.0, ax == 2 ? 1.0 : 0.0);
            double y[NV];
#pragma unroll
            for (int s = 0; s < NV; ++s) y[s] = (s < nv && on) ? jac_entry(s, anc0, anc1, point, dir) : 0.0;
            double sq = 0.0;
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                if (k < nv) {
                    y[k] = y[k] * sd[ldr + k];
                    sq += y[k] * y[k];
#pragma unroll
                    for (int s = k + 1; s < NV; ++s)
                        if (s < nv) y[s] -= sd[ld + s * nv + k] * y[k];
                }
            }
            sd[T.f + lane] = sq;
            wsync();
            if (on && ax == 0) {
                const int o = T.f + lane;
                sd[T.invw + i] = ((sd[o] + sd[o + 1]) + sd[o + 2]) / 3.0;
            }
            wsync();
        }
    }

    // y ← L⁻¹ y, column-oriented, with L held by rows across the lanes
    // (lane k: a[] = row k of L, rinv = 1/L_kk): each L entry reaches every
    // lane through v_readlane.
    __device__ __forceinline__ void fwd_regs(double (&y)[NV], const double (&a)[NV], double rinv) const {
        const int nv = m.nv;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if (k < nv) {
                y[k] = y[k] * readlane_f64(rinv, k);
#pragma unroll
                for (int s = k + 1; s < NV; ++s)
                    if (s < nv) y[s] = y[s] - readlane_f64(a[k], s) * y[k];
            }
        }
    }

    // invweights() with the factor in registers.
    __device__ __forceinline__ void invweights_regs(const double (&la)[NV], double lrinv, int nu) const {
        const int n = nu, nv = m.nv;
        // Passes of 63 lanes keep each link's three axes in one pass.
        for (int base = 0; base < 3 * n; base += 63) {
            const int idx = base + lane;
            const bool on = lane < 63 && idx < 3 * n;
            const int i = on ? si[T.lused + idx / 3] : 0, ax = idx % 3;
            const unsigned long long anc0 = tb.links[i].anc[0], anc1 = tb.links[i].anc[1];
            const V3<double> point = pos(i) + qrot(ori(i), z3<double>());
            const V3<double> dir = mk3(ax == 0 ? 1.0 : 0.0, ax == 1 ? 1.0 : 0.0, ax == 2 ? 1.0 : 0.0);
            double y[NV];
#pragma unroll
            for (int s = 0; s < NV; ++s) y[s] = (s < nv && on) ? jac_entry(s, anc0, anc1, point, dir) : 0.0;
            fwd_regs(y, la, lrinv);
            double sq = 0.0;
#pragma unroll
            for (int s = 0; s < NV; ++s)
                if (s < nv) sq += y[s] * y[s];
            sd[T.f + lane] = sq;
            wsync();
            if (on && ax == 0) {
                const int o = T.f + lane;
                sd[T.invw + i] = ((sd[o] + sd[o + 1]) + sd[o + 2]) / 3.0;
            }
            wsync();
        }
    }

    // Pyramidal rows n ± μ t of every active contact (solver.rs:2805-3120):
    // the Jacobian row in FP64 registers gives the bias and, through the
    // factor, the row y = L⁻¹ Jᵀ of Y, stored in R with the regularizer and
    // bias.  One lane per row.
    __device__ __forceinline__ void rows(int na, const RowStore<R>& rs, const double (&la)[NV], double lrinv) const {
        const int nv = m.nv, nr = 4 * na;
        const double dt = m.dt;
        for (int base = 0; base < nr; base += 64) {
            const int l = base + lane;
            const bool act = l < nr;
            const int c = si[T.alist + ((act ? l : 0) >> 2)], r = l & 3;
            const int o = T.cgeo + 9 * c;
            const GeomT& ga = tb.geoms[si[T.cids + 3 * c]];
            const GeomT& gb = tb.geoms[si[T.cids + 3 * c + 1]];
            const V3<double> cp = ld3(o), n = ld3(o + 3);
            const double pen = sd[o + 6] - sd[o + 8];
            const double mu = ga.friction < gb.friction ? ga.friction : gb.friction;
            // tangent_basis (world.rs:3104-3111)
            const V3<double> ref = nabs(n.z) < 0.9 ? mk3(0.0, 0.0, 1.0) : mk3(1.0, 0.0, 0.0);
            const V3<double> t1 = unit3(cross3(ref, n));
            const V3<double> t2 = cross3(n, t1);
            const V3<double> tt = r < 2 ? t1 : t2;
            const V3<double> dir = (r & 1) == 0 ? n + tt * mu : n - tt * mu;
            // Jacobian row (J·dir over both attachments), FP64 in registers.
            double y[NV];
#pragma unroll
            for (int s = 0; s < NV; ++s) y[s] = 0.0;
            double tr = 0.0;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const GeomT& g = side == 0 ? ga : gb;
                if (g.link >= 0) {
                    const double sign = side == 0 ? 1.0 : -1.0;
                    const V3<double> lp = pos(g.link);
                    const Qt<double> lo = ori(g.link);
                    const V3<double> point = lp + qrot(lo, qrot_inv(lo, cp - lp));
                    const unsigned long long anc0 = tb.links[g.link].anc[0], anc1 = tb.links[g.link].anc[1];
                    tr = tr + sd[T.invw + g.link];
#pragma unroll
                    for (int s = 0; s < NV; ++s)
                        if (s < nv) y[s] += jac_entry(s, anc0, anc1, point, dir) * sign;
                }
            }
            double vel = -0.0, fvel = -0.0;
#pragma unroll
            for (int s = 0; s < NV; ++s) {
                if (s < nv) {
                    vel = vel + y[s] * sd[T.qd + s];
                    fvel = fvel + y[s] * (sd[T.qfree + s] * dt);
                }
            }
            // SolRef / SolImp mixing and reference acceleration
            // (solver.rs:319-384, 3067-3120)
            double sr0 = ga.solref[0] <= gb.solref[0] ? ga.solref[0] : gb.solref[0];
            double sr1 = ga.solref[1] <= gb.solref[1] ? ga.solref[1] : gb.solref[1];
            if ((ga.solref[0] < 0.0 || gb.solref[0] < 0.0) && !(ga.solref[0] < 0.0 && gb.solref[0] < 0.0)) {
                const GeomT& gd = ga.solref[0] < 0.0 ? ga : gb;
                sr0 = gd.solref[0];
                sr1 = gd.solref[1];
            }
            double dmin = ga.solimp[0] < gb.solimp[0] ? ga.solimp[0] : gb.solimp[0];
            double dmax = ga.solimp[1] < gb.solimp[1] ? ga.solimp[1] : gb.solimp[1];
            double width = ga.solimp[2] < gb.solimp[2] ? ga.solimp[2] : gb.solimp[2];
            double mid = ga.solimp[3];
            int power = ga.power;
            const double lo = 1.0e-4, hi = 0.9999;
            auto clampi = [&](double v) { return v < lo ? lo : (v > hi ? hi : v); };
            dmin = clampi(dmin);
            dmax = clampi(dmax);
            mid = clampi(mid);
            width = width > 0.0 ? width : 0.0;
            if (power < 1) power = 1;
            auto impedance = [&](double position) {
                double x = (position - 0.0) / width;
                if (x < 0.0) x = -x;
                double y;
                if (x <= 0.0) {
                    y = 0.0;
                } else if (x >= 1.0) {
                    y = 1.0;
                } else if (x < mid) {
                    double b = x / mid, pw = 1.0;
                    for (int k2 = 0; k2 < power; ++k2) pw = pw * b;
                    y = mid * pw;
                } else {
                    double b = (1.0 - x) / (1.0 - mid), pw = 1.0;
                    for (int k2 = 0; k2 < power; ++k2) pw = pw * b;
                    y = 1.0 - (1.0 - mid) * pw;
                }
                return dmin + (dmax - dmin) * y;
            };
            const double position = -pen;
            const double imp = impedance(position);
            double tc = sr0;
            const double dr = sr1;
            const bool direct = tc < 0.0;
            if (!direct && tc < 2.0 * dt) tc = 2.0 * dt;
            double bcoef, kcoef;
            if (direct) {
                bcoef = -dr / dmax;
                kcoef = -tc / (dmax * dmax);
            } else {
                bcoef = 2.0 / (dmax * tc);
                kcoef = imp / (dmax * dmax * tc * tc * dr * dr);
            }
            const double diag_approx = tr * (1.0 + mu * mu);
            const double rbase = (1.0 - imp) / imp * diag_approx;
            const double refacc = -bcoef * vel - kcoef * position;
            if (act) {
                rs.reg[l] = 2.0 * mu * mu * rbase;
                rs.bias[l] = fvel - refacc * dt;
            }
            // Y row: L y = J_l (solver.rs:3045-3065 factor solves), stored in
            // R; the Delassus diagonal A_ll = ‖y‖² from the FP64 row.
            fwd_regs(y, la, lrinv);
            double dg = 0.0;
#pragma unroll
            for (int s = 0; s < NV; ++s)
                if (s < nv) dg += y[s] * y[s];
            if (act) {
                rs.diag[l] = dg;
#pragma unroll
                for (int s = 0; s < NV; ++s)
                    if (s < nv) rs.J[l * nv + s] = (R)y[s];
            }
        }
        wsync();
    }

    // PGS sweeps over NR rows (solver.rs:3269-3331).  Lane j holds row j: its
    // row of A (a), residual r_j = b_j + Σ A_jk f_k, impulse f_j.  Row i's
    // update f_i ← max(f_i − (r_i + R_i f_i)/(A_ii + R_i), 0) is written
    // max(c_i f_i − d_i r_i, 0) with d_i = 1/(A_ii + R_i), c_i = 1 − R_i d_i;
    // every lane evaluates it for its own row, lane i's result is broadcast,
    // kept by lane i, and every residual moves by A_ji Δf_i.  Rows
    // past the real count carry c = d = 0 and a zero row of A, so they stay
    // at zero and move nothing.  FP64 in every mode (FP32 residuals put the
    // mixed step's qacc over 1e-5 on box-box chains).
    template <int NR, class P>
    __device__ __forceinline__ void sweeps(const P* a, P cf, P invd, P& r, P& f) const {
        for (int it = 0; it < m.iterations; ++it) {
            const P g = cf * f;  // f_i changes only at row i
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                P fn = g - invd * r;
                fn = fn > P(0) ? fn : P(0);
                const P dl = readlane_r(fn - f, i);
                f = put_lane(f, fn, lane, i);
                r = r + a[i] * dl;
            }
        }
    }

    // Delassus + PGS (solver.rs:3045-3065, 3269-3331) for nr <= 64 rows; on
    // return sd[u] = Y f/dt = L⁻¹ Jᵀ f/dt.  A = Y Yᵀ in FP64 sums: on the
    // matrix cores up to 32 rows, on the vector units above.  The diagonal
    // comes from the FP64 row; the sweeps and u = Y f/dt run in FP64.
    // Row `lane` of A = Y Yᵀ for nr <= 16 NB rows, on the matrix cores:
    // v_mfma_f64_16x16x4f64 on 16 x 16 tiles (I, J) over the dofs four at a
    // time, FP64 sums of the stored rows (products of FP32 rows are exact).
    // Lane l feeds Y[16 I + l % 16][k0 + l / 16] as the first operand and
    // Y[16 J + l % 16][k0 + l / 16] as the second; register r of lane l then
    // holds A[16 I + 4 r + l / 16][16 J + l % 16].  A is symmetric, so lane
    // j = 16 J + m collects column j from the four lanes 16 p + m (p = 0..3):
    // one shuffle per (J, p, I, r), kept by the lanes of row block J.
    template <int NB>
    __device__ __forceinline__ void delassus_mfma(int nr, double (&a)[64]) const {
        typedef double f64x4 __attribute__((ext_vector_type(4)));
        f64x4 t[NB][NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int J = 0; J < NB; ++J) t[I][J] = f64x4{0.0, 0.0, 0.0, 0.0};
        const int nv = m.nv, lr = lane & 15, lk = lane >> 4;
        for (int k0 = 0; k0 < nv; k0 += 4) {
            const int k = k0 + lk;
            double y[NB];
#pragma unroll
            for (int I = 0; I < NB; ++I) {
                const int row = 16 * I + lr;
                y[I] = (row < nr && k < nv) ? (double)sJ[row * nv + k] : 0.0;
            }
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int J = 0; J < NB; ++J) t[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(y[I], y[J], t[I][J], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < 64; ++c) a[c] = 0.0;
#pragma unroll
        for (int J = 0; J < NB; ++J) {
            const bool keep = lk == J;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int src = 16 * p + lr;
#pragma unroll
                for (int I = 0; I < NB; ++I) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const double x = __shfl(t[I][J][r], src);
                        a[16 * I + 4 * r + p] = keep ? x : a[16 * I + 4 * r + p];
                    }
                }
            }
        }
    }

    __device__ __forceinline__ void pgs_small(int nr) const {
        const int nv = m.nv;
        const bool active = lane < nr;
        double a[64];
        if (nr <= 16) {
            delassus_mfma<1>(nr, a);
            pgs_rows(nr, a);
            return;
        }
        if (nr <= 32) {
            delassus_mfma<2>(nr, a);
            pgs_rows(nr, a);
            return;
        }
        // Row `lane` of A = Y Yᵀ, eight rows at a time.
        const R* yl = sJ + (active ? lane : 0) * nv;
#pragma unroll
        for (int ib = 0; ib < 64; ib += 8) {
            double acc[8];
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) acc[ii] = 0.0;
            if (ib < nr) {
                for (int k = 0; k < nv; ++k) {
                    const double yk = (double)yl[k];
#pragma unroll
                    for (int ii = 0; ii < 8; ++ii)
                        if (ib + ii < nr) acc[ii] += yk * (double)sJ[(ib + ii) * nv + k];
                }
            }
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) a[ib + ii] = (active && ib + ii < nr) ? acc[ii] : 0.0;
        }
        pgs_rows(nr, a);
    }

    // The sweeps on row `lane` of A, then u = Y f/dt.
    __device__ __forceinline__ void pgs_rows(int nr, const double (&a)[64]) const {
        const int nv = m.nv;
        const bool active = lane < nr;
        const double diag = active ? sd[T.rdg + lane] : 0.0;
        const double reg = active ? sd[T.rreg + lane] : 0.0;
        const double invd = active ? 1.0 / (diag + reg) : 0.0;
        const double cf = active ? 1.0 - reg * invd : 0.0;
        double r = active ? sd[T.rbias + lane] : 0.0;
        double f = 0.0;
        if (nr <= 16) sweeps<16>(a, cf, invd, r, f);
        else if (nr <= 32) sweeps<32>(a, cf, invd, r, f);
        else if (nr <= 48) sweeps<48>(a, cf, invd, r, f);
        else sweeps<64>(a, cf, invd, r, f);
        const double dt = m.dt;
        sd[T.f + lane] = active ? f / dt : 0.0;
        wsync();
        for (int s = lane; s < nv; s += 64) {
            double acc = 0.0;
            for (int j = 0; j < nr; ++j) acc += (double)sJ[j * nv + s] * sd[T.f + j];
            sd[T.u + s] = acc;
        }
        wsync();
    }

    // Same for nr > 64 rows (Y rows in the world's global overflow block,
    // per-row scalars in LDS), without forming A: residuals from z = Y f,
    // kept in lanes 0..nv-1 (r_i = b_i + y_i·z).
    __device__ __forceinline__ void pgs_large(int nr, const RowStore<R>& rs) const {
        const int nv = m.nv;
        for (int j = lane; j < nr; j += 64) {
            rs.diag[j] = rs.diag[j] + rs.reg[j];
            rs.f[j] = 0.0;
        }
        gsync();
        double z = 0.0;
        for (int it = 0; it < m.iterations; ++it) {
            for (int i = 0; i < nr; ++i) {
                const double yi = lane < nv ? (double)rs.J[i * nv + lane] : 0.0;
                const double dot = wave_sum(yi * z);
                const double fi = rs.f[i];
                double fn = fi - ((rs.bias[i] + dot) + rs.reg[i] * fi) / rs.diag[i];
                fn = fn > 0.0 ? fn : 0.0;
                z += yi * (fn - fi);
                rs.f[i] = fn;  // every lane stores the same value and reads back its own
            }
        }
        const double dt = m.dt;
        for (int s = lane; s < nv; s += 64) {
            double acc = 0.0;
            for (int j = 0; j < nr; ++j) acc += (double)rs.J[j * nv + s] * (rs.f[j] / dt);
            sd[T.u + s] = acc;
        }
        wsync();
    }

    // qacc = qfree + L⁻ᵀ u.
    __device__ __forceinline__ void back_subst_add() const {
        const double x = bwd_vec(lane < m.nv ? sd[T.u + lane] : 0.0);
        if (lane < m.nv) sd[T.qacc + lane] = sd[T.qfree + lane] + x;
        wsync();
    }

    // euler_step_with_workspace + integrate_tree_positions
    // (tree.rs:1407-1475).
    __device__ __forceinline__ void integrate() const {
        const double dt = m.dt;
        for (int s = lane; s < m.nv; s += 64) sd[T.qd + s] = sd[T.qd + s] + sd[T.qacc + s] * dt;
        wsync();
        for (int i = lane; i < m.n_links; i += 64) {
            const LinkT& l = tb.links[i];
            if (l.joint == JF_HINGE || l.joint == JF_SLIDE) {
                sd[T.q + l.q_off] = sd[T.q + l.q_off] + sd[T.qd + l.v_off] * dt;
            } else if (l.joint == JF_FREE) {
                const int qo = T.q + l.q_off, vo = T.qd + l.v_off;
                Qt<double> o{sd[qo + 3], sd[qo + 4], sd[qo + 5], sd[qo + 6]};
                V3<double> w = mk3(sd[vo], sd[vo + 1], sd[vo + 2]);
                V3<double> dp = qrot(o, mk3(sd[vo + 3], sd[vo + 4], sd[vo + 5])) * dt;
                sd[qo] = sd[qo] + dp.x;
                sd[qo + 1] = sd[qo + 1] + dp.y;
                sd[qo + 2] = sd[qo + 2] + dp.z;
                Qt<double> nx = qintegrate(o, w, dt);
                sd[qo + 3] = nx.x;
                sd[qo + 4] = nx.y;
                sd[qo + 5] = nx.z;
                sd[qo + 6] = nx.w;
            }
        }
        wsync();
    }

    // The limit rows of build_tree_solver_rows (solver.rs:2337-2373): one
    // per hinge/slide outside its range, in link order (coefficient +1 below
    // the range, -1 above).  Returns the count, uniform.
    __device__ __forceinline__ int limit_rows() const {
        int base = 0;
        for (int i0 = 0; i0 < m.n_links; i0 += 64) {
            const int i = i0 + lane;
            double c = 0.0, viol = 0.0;
            int slot = 0, li = -1;
            if (i < m.n_links && tb.links[i].lim >= 0) {
                const LinkT& l = tb.links[i];
                const LimitT& lt = tb.limits[l.lim];
                const double q = sd[T.q + l.q_off];
                if (q < lt.lo) {
                    c = 1.0;
                    viol = lt.lo - q;
                    li = l.lim;
                } else if (q > lt.hi) {
                    c = -1.0;
                    viol = q - lt.hi;
                    li = l.lim;
                }
                slot = l.v_off;
            }
            const unsigned long long bal = __ballot(li >= 0);
            if (li >= 0) {
                const int k = base + __popcll(bal & lanes_below(lane));
                si[T.lsl + k] = slot;
                si[T.llm + k] = li;
                sd[T.lsg + k] = c;
                sd[T.lvi + k] = viol;
            }
            base += __popcll(bal);
        }
        wsync();
        return __builtin_amdgcn_readfirstlane(base);
    }

    // solve_tree_limits (solver.rs:2509-2628) for nl rows: lane i holds row
    // i, y_i = L⁻¹(c_i e_slot) with L of M (no damping term) from sd[ld],
    // 1/L_kk from sd[ldr]; A = Y Yᵀ in FP64; the PGS update of sweeps()
    // (same projection, f >= 0) with R_i = (1-d)/d A_ii.  The per-dof force
    // goes to sd[T.qlim].
    __device__ __forceinline__ void limit_solve(int nl, int ld, int ldr) const {
        const int nv = m.nv;
        const double dt = m.dt;
        const bool act = lane < nl;
        double* y = sd + T.ylim + (act ? lane : 0) * nv;
        int slot = 0;
        double c = 0.0;
        if (act) {
            slot = si[T.lsl + lane];
            c = sd[T.lsg + lane];
            for (int s = 0; s < nv; ++s) y[s] = 0.0;
            y[slot] = 0.0 + c;
            for (int k = slot; k < nv; ++k) {
                const double yk = y[k] * sd[ldr + k];
                y[k] = yk;
                for (int s = k + 1; s < nv; ++s) y[s] -= sd[ld + s * nv + k] * yk;
            }
        }
        wsync();
        double a[64];
#pragma unroll
        for (int j = 0; j < 64; ++j) {
            double acc = 0.0;
            if (j < NV && act && j < nl) {
                const double* yj = sd + T.ylim + j * nv;
                for (int k = 0; k < nv; ++k) acc += y[k] * yj[k];
            }
            a[j] = acc;
        }
        double dg = 0.0, cf = 0.0, invd = 0.0, r = 0.0, f = 0.0;
        if (act) {
            for (int k = 0; k < nv; ++k) dg += y[k] * y[k];
            const LimitT& lt = tb.limits[si[T.llm + lane]];
            const SolImpE sim = solimp_effective(lt.solimp, lt.power);
            const double viol = sd[T.lvi + lane];
            const double vrow = 0.0 + c * sd[T.qd + slot];
            r = -reference_accel(-viol, vrow, lt.solref[0], lt.solref[1], sim, dt) * dt;
            const double d = impedance_at(sim, -viol);
            const double reg = d > 0.0 ? (1.0 - d) / d * dg : 0.0;
            invd = 1.0 / (dg + reg);
            cf = 1.0 - reg * invd;
        }
        if (nl <= 16) sweeps<16>(a, cf, invd, r, f);
        else if (NV <= 32 || nl <= 32) sweeps<32>(a, cf, invd, r, f);
        else if (nl <= 48) sweeps<48>(a, cf, invd, r, f);
        else sweeps<64>(a, cf, invd, r, f);
        for (int s = lane; s < nv; s += 64) sd[T.qlim + s] = 0.0;
        wsync();
        if (act) sd[T.qlim + slot] = 0.0 + c * (f / dt);
        wsync();
    }

    // The contact force recovered as a generalized force, Jᵀ f/dt = L·(Y f/dt)
    // with u = Y f/dt in sd[T.u] (solver.rs:3334-3364), added to the applied
    // force as step_trees does (world.rs:1899-1906).
    __device__ __forceinline__ void add_contact_force(bool with_limits) const {
        const int nv = m.nv;
        if (lane < nv) {
            double s = 0.0;
            for (int k = 0; k <= lane; ++k) s += sd[T.Li + lane * nv + k] * sd[T.u + k];
            sd[T.tau + lane] = sd[T.qfrc + lane] + ((with_limits ? sd[T.qlim + lane] : 0.0) + s);
        }
        wsync();
    }

    // rk4_step_with_workspace (tree.rs:1234-1364): sd[T.tau] held over four
    // explicit ABA stages (no implicit damping); qacc reports the first
    // stage.  Poses and xups describe the start state on entry.
    __device__ __forceinline__ void rk4() const {
        const double dt = m.dt;
        for (int k = lane; k < m.nq; k += 64) sd[T.q0 + k] = sd[T.q + k];
        for (int k = lane; k < m.nv; k += 64) sd[T.qd0 + k] = sd[T.qd + k];
        wsync();
        for (int st = 0; st < 4; ++st) {
            if (st > 0) {
                kinematics();
                store_xups();
            }
            const int out = st == 0 ? T.qacc : T.qfree;
            aba(out, false);
            const double h = st == 2 ? dt : dt * 0.5;  // advance_tree_state (tree.rs:1551-1573)
            // tree_deriv (tree.rs:1485-1540), lane per link; sums
            // ((k1 + 2 k2) + 2 k3) + k4 in stage order.
            for (int i = lane; i < m.n_links; i += 64) {
                const LinkT& l = tb.links[i];
                if (l.joint == JF_FIXED) continue;
                const int qo = l.q_off, vo = l.v_off;
                const int nqi = l.joint == JF_FREE ? 7 : 1, nvi = l.joint == JF_FREE ? 6 : 1;
                double dq[7];
                if (l.joint == JF_FREE) {
                    const Qt<double> o{sd[T.q + qo + 3], sd[T.q + qo + 4], sd[T.q + qo + 5], sd[T.q + qo + 6]};
                    const V3<double> vw = qrot(o, ld3(T.qd + vo + 3));
                    const Qt<double> dqo = qderiv(o, ld3(T.qd + vo));
                    dq[0] = vw.x;
                    dq[1] = vw.y;
                    dq[2] = vw.z;
                    dq[3] = dqo.x;
                    dq[4] = dqo.y;
                    dq[5] = dqo.z;
                    dq[6] = dqo.w;
                } else {
                    dq[0] = sd[T.qd + vo];
                }
                for (int k = 0; k < nqi; ++k) {
                    const int a = qo + k;
                    sd[T.aq + a] = st == 0 ? dq[k] : (st == 3 ? sd[T.aq + a] + dq[k] : sd[T.aq + a] + 2.0 * dq[k]);
                    if (st < 3) sd[T.q + a] = sd[T.q0 + a] + dq[k] * h;
                }
                for (int k = 0; k < nvi; ++k) {
                    const int a = vo + k;
                    const double dv = sd[out + a];
                    sd[T.av + a] = st == 0 ? dv : (st == 3 ? sd[T.av + a] + dv : sd[T.av + a] + 2.0 * dv);
                    if (st < 3) sd[T.qd + a] = sd[T.qd0 + a] + dv * h;
                }
            }
            wsync();
        }
        const double w6 = dt * (1.0 / 6.0);
        for (int k = lane; k < m.nq; k += 64) sd[T.q + k] = sd[T.q0 + k] + sd[T.aq + k] * w6;
        for (int k = lane; k < m.nv; k += 64) sd[T.qd + k] = sd[T.qd0 + k] + sd[T.av + k] * w6;
        wsync();
        if (lane == 0 && tb.links[0].joint == JF_FREE) {
            const int qo = T.q + tb.links[0].q_off;
            const Qt<double> o = qnormalize(Qt<double>{sd[qo + 3], sd[qo + 4], sd[qo + 5], sd[qo + 6]});
            sd[qo + 3] = o.x;
            sd[qo + 4] = o.y;
            sd[qo + 5] = o.z;
            sd[qo + 6] = o.w;
        }
        wsync();
    }

    // One World::step (world.rs:1459-1537); poses and link velocities
    // describe the current state on entry and on exit.  Returns a fault code.
    __device__ __forceinline__ int step(int* np, int* ncon) const {
        const int nc = contacts();
        if (nc > m.cap_c) return 8;  // NEWT_ERR_CAPACITY
        const int na = m.mode == 1 ? active_contacts(nc) : 0;
        const bool rk4_ = X && m.integrator == 2;
        store_xups();
        for (int s = lane; s < m.nv; s += 64) sd[T.tau + s] = sd[T.qfrc + s];
        wsync();
        // Limit rows at the start state (step_trees_euler / step_trees,
        // world.rs:1826-2051); PGS mode only.
        int nl = 0;
        if constexpr (X) nl = m.n_ranged > 0 && m.mode == 1 ? limit_rows() : 0;
        if (na > 0 || nl > 0) {
            mass_matrix();
            double la[NV], lrinv;
            if (m.damped) {
                if (!factor(T.Ld, T.Ldr, la, lrinv)) {
                    if (na > 0) return 3;  // NEWT_ERR_NOT_PD
                    nl = 0;                // no limit force (solver.rs:2529-2532)
                } else {
                    if (na > 0) invweights(T.Ld, T.Ldr, contact_links(na));
                    if constexpr (X)
                        if (nl > 0) limit_solve(nl, T.Ld, T.Ldr);
                }
                add_damping();
            }
            if (!factor(T.Li, T.Lr, la, lrinv)) {
                if (na > 0) return 3;
                nl = 0;
            } else if (!m.damped) {
                if (na > 0) invweights_regs(la, lrinv, contact_links(na));
                if constexpr (X)
                    if (nl > 0) limit_solve(nl, T.Li, T.Lr);
            }
            if (na > 0) {
                // free acceleration: (M + dt·D) q̈ = τ − D q̇ − C
                rne_rhs(T.qfree);
                {
                    const double qf = bwd_vec(fwd_vec(lane < m.nv ? sd[T.qfree + lane] : 0.0, la, lrinv));
                    if (lane < m.nv) sd[T.qfree + lane] = qf;
                    wsync();
                }
                if (4 * na <= 64) {
                    rows(na, RowStore<R>{sJ, sd + T.rbias, sd + T.rreg, sd + T.rdg, sd + T.f}, la, lrinv);
                    pgs_small(4 * na);
                } else {
                    rows(na, ov, la, lrinv);
                    gsync();
                    pgs_large(4 * na, ov);
                }
                if (rk4_) {
                    if constexpr (X) add_contact_force(nl > 0);
                } else {
                    if (X && nl > 0) {
                        // u += L⁻¹ τ_limit: q̈ = q̈_free + (M + dt·D)⁻¹(Jᵀf/dt + τ_limit)
                        const double x = fwd_vec(lane < m.nv ? sd[T.qlim + lane] : 0.0, la, lrinv);
                        if (lane < m.nv) sd[T.u + lane] = sd[T.u + lane] + x;
                        wsync();
                    }
                    back_subst_add();
                }
            } else {
                if (nl > 0) {
                    for (int s = lane; s < m.nv; s += 64) sd[T.tau + s] = sd[T.qfrc + s] + (sd[T.qlim + s] + 0.0);
                    wsync();
                }
                if (!rk4_) aba(T.qacc, true);
            }
        } else if (!rk4_) {
            aba(T.qacc, true);
        }
        if constexpr (X) {
            if (rk4_)
                rk4();
            else
                integrate();
        } else {
            integrate();
        }
        kinematics();
        populate(false);
        *np = __builtin_amdgcn_readfirstlane(si[T.misc]);
        *ncon = nc;
        return 0;
    }
};

// Largest workgroup of the persistent step, in worlds (wavefronts).
template <int NV> constexpr int team_max_worlds() { return NV <= 24 ? 8 : 4; }

// One batch of worlds sharing a model, as the persistent step reads it: the
// model's scalars and LDS table layout, the per-world LDS layout, the state
// and output arrays, and the workgroups it occupies in the grid.
struct TeamJob {
    ModelT m;
    TeamLayout T;
    CacheLayout C;
    const int4* blob;  // model tables in the layout of C
    Layout L;          // offsets of the contact and pair outputs in dws / iws
    DataT d;
    double* dws;
    int* iws;
    char* ovf;         // per-world blocks for contact sets over 64 rows
    int wpb;           // worlds (wavefronts) per workgroup
    int first_block;   // first workgroup of this job in the grid
};

// Persistent step kernel over a table of jobs (sorted by first_block; one
// job for a newt_data, several for a batch of datas of different models).
// Workgroup b runs the last job whose first_block <= b: its `wpb`
// wavefronts step worlds (b - first_block)*wpb .. +wpb-1, one wavefront
// each, n_steps times.  The model tables are copied once into the
// workgroup's LDS (shared by its worlds); each world's state follows at
// C.bytes + wave*T.bytes.  Global state is lane-interleaved ([element][W]).
// At NV <= 24 the register budget is held to 256 so two wavefronts fit a
// SIMD (eight per CU).
template <class R, int NV, bool X>
__global__ void __launch_bounds__(64 * team_max_worlds<NV>())
    __attribute__((amdgpu_waves_per_eu(NV <= 24 ? 2 : 1)))
    k_team(const TeamJob* __restrict__ jobs, int n_jobs, int n_steps) {
    extern __shared__ __align__(16) double smem[];
    // The job of this workgroup: uniform, found by bisection.
    int lo = 0, hi = n_jobs;
    const int b = (int)blockIdx.x;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (jobs[mid].first_block <= b) lo = mid;
        else hi = mid;
    }
    const TeamJob& J = jobs[lo];
    const ModelT& m = J.m;
    const TeamLayout& T = J.T;
    const CacheLayout& C = J.C;
    const Layout& L = J.L;
    const DataT& d = J.d;
    double* __restrict__ dws = J.dws;
    int* __restrict__ iws = J.iws;
    const int wpb = J.wpb;
    char* lds = (char*)smem;
    for (int k = threadIdx.x; k < C.bytes / 16; k += blockDim.x) ((int4*)lds)[k] = J.blob[k];
    __syncthreads();
    const Tables tb{(const LinkT*)(lds + C.links),     (const GeomT*)(lds + C.geoms),
                    (const PairT*)(lds + C.pairs),     (const int*)(lds + C.level_off),
                    (const int*)(lds + C.level_links), (const int*)(lds + C.level_nrank),
                    (const int*)(lds + C.dof_link),    (const LimitT*)(lds + C.limits)};
    // The wavefront index is uniform within a wavefront; readfirstlane lets
    // the compiler treat everything derived from it as scalar.
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, W = d.W;
    const int w = (b - J.first_block) * wpb + wave;
    if (wave >= wpb || w >= W || d.fault[w] != 0) return;
    char* base = lds + C.bytes + wave * T.bytes;
    double* sd = (double*)base;
    int* si = (int*)(base + T.i_off);
    R* sJ = (R*)(base + T.j_off);
    const int cr = m.cap_r;
    const RowStore<R> ov{(R*)(J.ovf + (size_t)w * T.ov_stride), sd + T.lg, sd + T.lg + cr, sd + T.lg + 2 * cr,
                         sd + T.lg + 3 * cr};
    Team<R, NV, X> t{m, T, tb, lane, sd, si, sJ, ov};
    for (int k = lane; k < m.nq; k += 64) sd[T.q + k] = d.qpos[(size_t)k * W + w];
    for (int k = lane; k < m.nv; k += 64) {
        sd[T.qd + k] = d.qvel[(size_t)k * W + w];
        sd[T.qfrc + k] = d.qfrc[(size_t)k * W + w];
        sd[T.qacc + k] = d.qacc[(size_t)k * W + w];
    }
    for (int k = lane; k < 6 * m.n_geoms; k += 64) sd[T.fat + k] = d.fat[(size_t)k * W + w];
    for (int i = lane; i < m.n_links; i += 64) {
        const LinkT& l = tb.links[i];
        if (l.joint == JF_FREE)
            for (int k = 0; k < 6; ++k) sd[T.dvec + l.v_off + k] = l.free_damping;
        else if (l.joint == JF_HINGE || l.joint == JF_SLIDE)
            sd[T.dvec + l.v_off] = l.damping + 0.0;
    }
    wsync();
    t.kinematics();
    // update_broadphase at step start (world.rs:1190-1226): a no-op unless
    // the state was set since the last step; with init_bounds the proxies
    // are rebuilt from the state, as after a state setter.
    t.populate(d.init_bounds != 0);
    // The previous launch's counts, clamped to the compiled capacities (they
    // bound the copy-out below when no step runs).
    int fault = 0, np = d.init_bounds ? 0 : min(max(d.npair[w], 0), m.n_pairs),
        nc = d.init_bounds ? 0 : min(max(d.ncontact[w], 0), m.cap_c);
    for (int s = 0; s < n_steps && fault == 0; ++s) fault = t.step(&np, &nc);
    for (int k = lane; k < m.nq; k += 64) d.qpos[(size_t)k * W + w] = sd[T.q + k];
    for (int k = lane; k < m.nv; k += 64) {
        d.qvel[(size_t)k * W + w] = sd[T.qd + k];
        d.qacc[(size_t)k * W + w] = sd[T.qacc + k];
    }
    for (int k = lane; k < 6 * m.n_geoms; k += 64) d.fat[(size_t)k * W + w] = sd[T.fat + k];
    for (int k = lane; k < 9 * nc; k += 64) dws[(size_t)(L.cgeo + k) * W + w] = sd[T.cgeo + k];
    for (int k = lane; k < 3 * nc; k += 64) iws[(size_t)(L.cids + k) * W + w] = si[T.cids + k];
    for (int k = lane; k < np; k += 64) iws[(size_t)(L.pairs + k) * W + w] = si[T.plist + k];
    if (lane == 0) {
        d.npair[w] = np;
        d.ncontact[w] = nc;
        if (fault) d.fault[w] = fault;
    }
}

// Launch / LDS-limit helpers of one instantiation.  Each is instantiated in
// one of the newt_team*.hip units (so the units compile in parallel) and
// declared extern below for the dispatcher.
template <class R, int NV, bool X>
hipError_t launch_team(dim3 grid, dim3 block, int lds_bytes, hipStream_t stream, const TeamJob* jobs, int n_jobs,
                       int n_steps) {
    hipLaunchKernelGGL((k_team<R, NV, X>), grid, block, lds_bytes, stream, jobs, n_jobs, n_steps);
    return hipGetLastError();
}
template <class R, int NV, bool X>
hipError_t set_team_lds(int bytes) {
    return hipFuncSetAttribute((const void*)k_team<R, NV, X>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

#define NEWT_TEAM_EXTERN(R, NV, X)                                                                           \
    extern template hipError_t launch_team<R, NV, X>(dim3, dim3, int, hipStream_t, const TeamJob*, int, int); \
    extern template hipError_t set_team_lds<R, NV, X>(int);
#define NEWT_TEAM_INSTANCE(R, NV, X)                                                                  \
    template hipError_t launch_team<R, NV, X>(dim3, dim3, int, hipStream_t, const TeamJob*, int, int); \
    template hipError_t set_team_lds<R, NV, X>(int);

}  // namespace newt
