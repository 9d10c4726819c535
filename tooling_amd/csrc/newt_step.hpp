// newt_step.hpp — the per-world step executed by one lane of the gfx950 kernel.
//
// One lane steps one world.  Worlds are independent (nothing in
// World::step couples two worlds), so a wavefront advances 64 worlds in
// lockstep and all per-world arrays are lane-interleaved (element i of world
// w at base[i*W + w]): every load and store of a wavefront touches 64
// consecutive elements.  Model tables are uniform across the lanes and are
// read through the scalar cache.
//
// Phase order per step follows World::step (newt/src/world.rs:1459-1537)
// with Euler integration and the PGS constraint solver:
//   broadphase pairs -> narrowphase contacts -> tree contact solve
//   (CRB, Cholesky, free-velocity ABA, Jacobian rows, Delassus, PGS)
//   -> implicit-damping ABA -> semi-implicit Euler -> broadphase refresh.
// Two arithmetic types: R for the bulk (kinematics, ABA, rows, Delassus,
// PGS) and D = double for the mass-matrix factorization and its solves.
// The FP64 step instantiates R = double; the mixed step R = float.
#pragma once

#include "newt_math.hpp"
#include "newt_types.hpp"

namespace newt {

template <class T> struct Lane {
    T* p;
    int W, w;
    __device__ __forceinline__ T& operator[](int i) const { return p[(size_t)i * W + w]; }
};

// Rust's float `Sum`: a fold that starts at -0.0.
template <class R> struct Sum {
    R acc;
    __device__ __forceinline__ Sum() : acc(R(-0.0)) {}
    __device__ __forceinline__ void add(R x) { acc = acc + x; }
};

template <class R> struct World {
    const ModelT& m;
    const Layout& L;
    Lane<R> ws;
    Lane<double> dw;
    Lane<int> iw;
    Lane<double> qpos, qvel, qacc, qfrc, fat;

    // ---------------- state access -----------------------------------------
    __device__ R q(int k) const { return (R)qpos[k]; }
    __device__ R qd(int k) const { return (R)qvel[k]; }

    __device__ V3<R> pos(int i) const { return {ws[L.poses + i * 7], ws[L.poses + i * 7 + 1], ws[L.poses + i * 7 + 2]}; }
    __device__ Qt<R> ori(int i) const {
        return {ws[L.poses + i * 7 + 3], ws[L.poses + i * 7 + 4], ws[L.poses + i * 7 + 5], ws[L.poses + i * 7 + 6]};
    }
    __device__ void set_pose(int i, V3<R> p, Qt<R> o) const {
        ws[L.poses + i * 7] = p.x; ws[L.poses + i * 7 + 1] = p.y; ws[L.poses + i * 7 + 2] = p.z;
        ws[L.poses + i * 7 + 3] = o.x; ws[L.poses + i * 7 + 4] = o.y; ws[L.poses + i * 7 + 5] = o.z;
        ws[L.poses + i * 7 + 6] = o.w;
    }
    __device__ V3<R> ld3(int off) const { return {ws[off], ws[off + 1], ws[off + 2]}; }
    __device__ void st3(int off, V3<R> v) const { ws[off] = v.x; ws[off + 1] = v.y; ws[off + 2] = v.z; }
    __device__ SM<R> ldm(int off) const { return {ld3(off), ld3(off + 3)}; }
    __device__ void stm(int off, SM<R> s) const { st3(off, s.w); st3(off + 3, s.v); }
    __device__ SF<R> ldf(int off) const { return {ld3(off), ld3(off + 3)}; }
    __device__ void stf(int off, SF<R> s) const { st3(off, s.t); st3(off + 3, s.f); }
    __device__ Xf<R> ldx(int i) const {
        Xf<R> x;
#pragma unroll
        for (int k = 0; k < 9; ++k) x.E.m[k] = ws[L.xup + i * 12 + k];
        x.r = ld3(L.xup + i * 12 + 9);
        return x;
    }
    __device__ void stx(int i, const Xf<R>& x) const {
#pragma unroll
        for (int k = 0; k < 9; ++k) ws[L.xup + i * 12 + k] = x.E.m[k];
        st3(L.xup + i * 12 + 9, x.r);
    }

    // ---------------- kinematics (tree.rs:771-858, world.rs:2777-2854) -----
    __device__ void forward_kinematics() const {
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            V3<R> p;
            Qt<R> o;
            V3<R> pj = mk3((R)l.pj[0], (R)l.pj[1], (R)l.pj[2]);
            V3<R> cj = mk3((R)l.cj[0], (R)l.cj[1], (R)l.cj[2]);
            if (l.joint == JF_FREE) {
                int qo = l.q_off;
                p = mk3(q(qo), q(qo + 1), q(qo + 2));
                o = {q(qo + 3), q(qo + 4), q(qo + 5), q(qo + 6)};
            } else if (l.joint == JF_FIXED) {
                V3<R> pp = z3<R>();
                Qt<R> po = qid<R>();
                if (l.parent >= 0) { pp = pos(l.parent); po = ori(l.parent); }
                Qt<R> pjq{(R)l.pjq[0], (R)l.pjq[1], (R)l.pjq[2], (R)l.pjq[3]};
                Qt<R> cjq{(R)l.cjq[0], (R)l.cjq[1], (R)l.cjq[2], (R)l.cjq[3]};
                V3<R> jp = pp + qrot(po, pj);
                Qt<R> jo = qmul(po, pjq);
                o = qmul(jo, qconj(cjq));
                p = jp - qrot(o, cj);
            } else {
                V3<R> pp = pos(l.parent);
                Qt<R> po = ori(l.parent);
                V3<R> ax = mk3((R)l.axis[0], (R)l.axis[1], (R)l.axis[2]);
                R qq = q(l.q_off);
                V3<R> jp = pp + qrot(po, pj);
                if (l.joint == JF_HINGE) {
                    o = qmul(po, qaxis(ax, qq));
                    p = jp - qrot(o, cj);
                } else {
                    o = po;
                    p = (jp + qrot(po, ax * qq)) - qrot(o, cj);
                }
            }
            set_pose(i, p, o);
        }
    }

    // (linear velocity at COM, angular velocity), both world frame.
    __device__ void link_velocities() const {
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            V3<R> lin = z3<R>(), ang = z3<R>();
            if (l.parent < 0) {
                if (l.joint == JF_FREE) {
                    Qt<R> o = ori(i);
                    int vo = l.v_off;
                    ang = qrot(o, mk3(qd(vo), qd(vo + 1), qd(vo + 2)));
                    lin = qrot(o, mk3(qd(vo + 3), qd(vo + 4), qd(vo + 5)));
                }
            } else {
                V3<R> pv = ld3(L.lvel + l.parent * 6), pw = ld3(L.lvel + l.parent * 6 + 3);
                V3<R> cp = pos(i), pp = pos(l.parent);
                Qt<R> po = ori(l.parent);
                V3<R> pac = pv + cross3(pw, cp - pp);
                if (l.joint == JF_HINGE) {
                    V3<R> aw = qrot(po, mk3((R)l.axis[0], (R)l.axis[1], (R)l.axis[2]));
                    R r = qd(l.v_off);
                    ang = pw + aw * r;
                    V3<R> jw = pp + qrot(po, mk3((R)l.pj[0], (R)l.pj[1], (R)l.pj[2]));
                    lin = pac + cross3(aw * r, cp - jw);
                } else if (l.joint == JF_SLIDE) {
                    V3<R> aw = qrot(po, mk3((R)l.axis[0], (R)l.axis[1], (R)l.axis[2]));
                    lin = pac + aw * qd(l.v_off);
                    ang = pw;
                } else {  // fixed
                    lin = pac;
                    ang = pw;
                }
            }
            st3(L.lvel + i * 6, lin);
            st3(L.lvel + i * 6 + 3, ang);
        }
    }

    // geom_world_pose (geom.rs:887-898)
    __device__ void geom_pose(const GeomT& g, V3<R>* p, Qt<R>* o) const {
        V3<R> gp = mk3((R)g.pos[0], (R)g.pos[1], (R)g.pos[2]);
        Qt<R> gq{(R)g.quat[0], (R)g.quat[1], (R)g.quat[2], (R)g.quat[3]};
        if (g.link < 0) { *p = gp; *o = gq; return; }
        Qt<R> lo = ori(g.link);
        *p = pos(g.link) + qrot(lo, gp);
        *o = qmul(lo, gq);
    }

    // geom_aabb (broadphase.rs:708-770) as (min, max).
    __device__ void geom_aabb(const GeomT& g, V3<R> p, Qt<R> o, V3<R>* mn, V3<R>* mx) const {
        if (g.shape == SH_PLANE) {
            R inf = R(__builtin_inf());
            *mn = mk3(-inf, -inf, -inf);
            *mx = mk3(inf, inf, inf);
            return;
        }
        V3<R> e;
        if (g.shape == SH_SPHERE) {
            e = mk3((R)g.size[0], (R)g.size[0], (R)g.size[0]);
        } else if (g.shape == SH_BOX) {
            V3<R> x = abs3(qrot(o, mk3(R(1), R(0), R(0)))) * (R)g.size[0];
            V3<R> y = abs3(qrot(o, mk3(R(0), R(1), R(0)))) * (R)g.size[1];
            V3<R> z = abs3(qrot(o, mk3(R(0), R(0), R(1)))) * (R)g.size[2];
            e = (x + y) + z;
        } else {
            R r = (R)g.size[0];
            e = abs3(qrot(o, mk3(R(0), R(0), R(1)))) * (R)g.size[1] + mk3(r, r, r);
        }
        *mn = p - e;
        *mx = p + e;
    }

    // populate_broadphase_tree + swept_geom_bound + update_with_filter_and_tree
    // (world.rs:1254-1358, broadphase.rs:46-54, 320-346).  Needs current poses
    // and link velocities.
    __device__ void populate(R dt, bool insert_all) const {
        for (int gi = 0; gi < m.n_geoms; ++gi) {
            const GeomT& g = m.geoms[gi];
            V3<R> sp;
            Qt<R> so;
            geom_pose(g, &sp, &so);
            V3<R> smn, smx, emn, emx;
            geom_aabb(g, sp, so, &smn, &smx);
            R ang_speed = R(0.0);
            V3<R> pivot = sp;
            if (g.link >= 0) {
                V3<R> v = ld3(L.lvel + g.link * 6), w = ld3(L.lvel + g.link * 6 + 3);
                V3<R> lp = pos(g.link);
                Qt<R> lo = ori(g.link);
                V3<R> ep = lp + v * dt;
                Qt<R> eo = qintegrate(lo, qrot_inv(lo, w), dt);
                V3<R> gp = mk3((R)g.pos[0], (R)g.pos[1], (R)g.pos[2]);
                Qt<R> gq{(R)g.quat[0], (R)g.quat[1], (R)g.quat[2], (R)g.quat[3]};
                geom_aabb(g, ep + qrot(eo, gp), qmul(eo, gq), &emn, &emx);
                ang_speed = len3(w);
                pivot = lp;
            } else {
                emn = smn;
                emx = smx;
            }
            V3<R> bc = (smn + smx) * R(0.5);
            R br = len3((smx - smn) * R(0.5));
            R pr = len3(bc - pivot) + br;
            R sweep = (ang_speed > R(0.0) && nfinite(pr)) ? pr * ang_speed * dt : R(0.0);
            V3<R> mn = mk3(nmin(smn.x, emn.x), nmin(smn.y, emn.y), nmin(smn.z, emn.z));
            V3<R> mx = mk3(nmax(smx.x, emx.x), nmax(smx.y, emx.y), nmax(smx.z, emx.z));
            V3<R> s3 = mk3(sweep, sweep, sweep);
            mn = mn - s3;
            mx = mx + s3;
            R mg = (R)g.margin;
            V3<R> m3 = mk3(mg, mg, mg);
            mn = mn - m3;
            mx = mx + m3;
            int f0 = gi * 6;
            R fmnx = (R)fat[f0], fmny = (R)fat[f0 + 1], fmnz = (R)fat[f0 + 2];
            R fmxx = (R)fat[f0 + 3], fmxy = (R)fat[f0 + 4], fmxz = (R)fat[f0 + 5];
            bool inside = fmnx <= mn.x && fmny <= mn.y && fmnz <= mn.z && fmxx >= mx.x && fmxy >= mx.y && fmxz >= mx.z;
            if (!inside || insert_all) {
                V3<R> e = mx - mn;
                V3<R> fm = mk3(nmax(e.x * R(0.1), R(1.0e-4)), nmax(e.y * R(0.1), R(1.0e-4)), nmax(e.z * R(0.1), R(1.0e-4)));
                V3<R> a = mn - fm, b = mx + fm;
                fat[f0] = a.x; fat[f0 + 1] = a.y; fat[f0 + 2] = a.z;
                fat[f0 + 3] = b.x; fat[f0 + 4] = b.y; fat[f0 + 5] = b.z;
            }
        }
    }

    // ---------------- broadphase pairs + narrowphase -----------------------
    __device__ int active_pairs() const {
        int n = 0;
        for (int k = 0; k < m.n_pairs; ++k) {
            int a = m.pairs[k].a * 6, b = m.pairs[k].b * 6;
            bool ov = fat[a] <= fat[b + 3] && fat[a + 3] >= fat[b] && fat[a + 1] <= fat[b + 4] &&
                      fat[a + 4] >= fat[b + 1] && fat[a + 2] <= fat[b + 5] && fat[a + 5] >= fat[b + 2];
            if (ov) iw[L.pairs + n++] = k;
        }
        return n;
    }

    __device__ void emit(int* nc, int ga, int gb, int slot, V3<R> p, V3<R> n, R pen, R fr, R gap) const {
        int k = *nc;
        iw[L.cids + 3 * k] = ga;
        iw[L.cids + 3 * k + 1] = gb;
        iw[L.cids + 3 * k + 2] = slot;
        int o = L.cgeo + 9 * k;
        st3(o, p);
        st3(o + 3, n);
        ws[o + 6] = pen;
        ws[o + 7] = fr;
        ws[o + 8] = gap;
        *nc = k + 1;
    }

    // One primitive with `prim` the non-plane geom (contact.rs:269-403);
    // `flip` relabels A/B and negates the normal (contact.rs:2579-2588).
    __device__ void plane_primitive(int* nc, int ip, const GeomT& gp, int iq, const GeomT& gq, bool flip, R fr,
                                    R margin, R gap) const {
        V3<R> pp, qp;
        Qt<R> po, qo;
        geom_pose(gq, &pp, &po);  // plane pose
        geom_pose(gp, &qp, &qo);  // primitive pose
        V3<R> n = qrot(po, mk3(R(0), R(0), R(1)));
        int slot = 0;
        int ga = flip ? iq : ip, gb = flip ? ip : iq;
        V3<R> nn = flip ? -n : n;
        if (gp.shape == SH_SPHERE) {
            R radius = (R)gp.size[0];
            R raw = dot3(qp - pp, n) - radius;
            if (raw <= margin) emit(nc, ga, gb, slot++, qp - n * (radius + raw * R(0.5)), nn, margin - raw, fr, gap);
        } else if (gp.shape == SH_BOX) {
            R dist = dot3(qp - pp, n);
            for (int i = 0; i < 8; ++i) {
                V3<R> loc = mk3((i & 1) == 0 ? -(R)gp.size[0] : (R)gp.size[0], (i & 2) == 0 ? -(R)gp.size[1] : (R)gp.size[1],
                                (i & 4) == 0 ? -(R)gp.size[2] : (R)gp.size[2]);
                V3<R> co = qrot(qo, loc);
                R ld = dot3(n, co);
                R raw = dist + ld;
                if (raw > margin || ld > R(0.0)) continue;
                if (slot >= 4) continue;
                emit(nc, ga, gb, slot++, (qp + co) - n * (raw * R(0.5)), nn, margin - raw, fr, gap);
            }
        } else if (gp.shape == SH_CAPSULE) {
            R radius = (R)gp.size[0], hh = (R)gp.size[1];
            V3<R> aw = qrot(qo, mk3(R(0), R(0), R(1)));
            for (int e = 0; e < 2; ++e) {
                V3<R> end = e == 0 ? qp + aw * hh : qp - aw * hh;
                R raw = dot3(end - pp, n) - radius;
                if (raw <= margin) emit(nc, ga, gb, slot++, end - n * (radius + raw * R(0.5)), nn, margin - raw, fr, gap);
            }
        }
    }

    __device__ int narrowphase(int npairs) const {
        int nc = 0;
        for (int k = 0; k < npairs; ++k) {
            const PairT& pr = m.pairs[iw[L.pairs + k]];
            const GeomT& ga = m.geoms[pr.a];
            const GeomT& gb = m.geoms[pr.b];
            R fr = (R)ga.friction < (R)gb.friction ? (R)ga.friction : (R)gb.friction;
            R margin = (R)ga.margin > (R)gb.margin ? (R)ga.margin : (R)gb.margin;
            R gap = (R)ga.gap > (R)gb.gap ? (R)ga.gap : (R)gb.gap;
            // (primitive, plane) in the given order, else swapped.
            if (gb.shape == SH_PLANE) plane_primitive(&nc, pr.a, ga, pr.b, gb, false, fr, margin, gap);
            else plane_primitive(&nc, pr.b, gb, pr.a, ga, true, fr, margin, gap);
        }
        return nc;
    }

    // ---------------- Xup, CRB, Cholesky -----------------------------------
    // aba.rs transform (762-791); crb=true uses xup_for_link_hinge's -q
    // (tree.rs:1134-1140) instead of (0 - q): they differ only in the sign
    // of a zero.
    __device__ Xf<R> xup(int i, bool crb) const {
        const LinkT& l = m.links[i];
        Xf<R> x;
        if (l.joint == JF_FIXED) {
#pragma unroll
            for (int k = 0; k < 9; ++k) x.E.m[k] = (R)l.xfE[k];
            x.r = mk3((R)l.xfr[0], (R)l.xfr[1], (R)l.xfr[2]);
        } else if (l.joint == JF_HINGE) {
            V3<R> ax = unit3(mk3((R)l.axis[0], (R)l.axis[1], (R)l.axis[2]));
            R qq = q(l.q_off);
            R s, c;
            nsincos(crb ? (-qq) * R(0.5) : (R(0) - qq) * R(0.5), &s, &c);
            x.E = qmat(Qt<R>{ax.x * s, ax.y * s, ax.z * s, c});
            x.r = mk3((R)l.cj[0], (R)l.cj[1], (R)l.cj[2]) - mv(x.E, mk3((R)l.pj[0], (R)l.pj[1], (R)l.pj[2]));
        } else if (l.joint == JF_SLIDE) {
            for (int k = 0; k < 9; ++k) x.E.m[k] = (k % 4 == 0) ? R(1) : R(0);
            x.r = (mk3((R)l.cj[0], (R)l.cj[1], (R)l.cj[2]) - mk3((R)l.pj[0], (R)l.pj[1], (R)l.pj[2])) -
                  mk3((R)l.axis[0], (R)l.axis[1], (R)l.axis[2]) * q(l.q_off);
        } else {
            for (int k = 0; k < 9; ++k) x.E.m[k] = (k % 4 == 0) ? R(1) : R(0);
            x.r = z3<R>();
        }
        return x;
    }
    __device__ SM<R> subspace(int i) const {
        const LinkT& l = m.links[i];
        V3<R> ax = mk3((R)l.axis[0], (R)l.axis[1], (R)l.axis[2]);
        if (l.joint == JF_HINGE) return {ax, cross3(mk3((R)l.cj[0], (R)l.cj[1], (R)l.cj[2]), ax)};
        return {z3<R>(), ax};
    }

    // dynamics::mass_matrix (dynamics.rs:474-578) into dw[M] (row-major nv*nv).
    __device__ void mass_matrix() const {
        const int nv = m.nv, n = m.n_links;
        for (int k = 0; k < nv * nv; ++k) dw[L.M + k] = 0.0;
        for (int i = 0; i < n; ++i) {
            for (int k = 0; k < 36; ++k) ws[L.ia + i * 36 + k] = (R)m.links[i].si[k];
            stx(i, xup(i, true));
        }
        for (int i = n - 1; i >= 1; --i) {
            int p = m.links[i].parent;
            R a[36], out[36];
            for (int k = 0; k < 36; ++k) a[k] = ws[L.ia + i * 36 + k];
            m6pullback(a, ldx(i), out);
            for (int k = 0; k < 36; ++k) ws[L.ia + p * 36 + k] = ws[L.ia + p * 36 + k] + out[k];
        }
        if (m.links[0].joint == JF_FREE) {
            int o = m.links[0].v_off;
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) dw[L.M + (o + r) * nv + (o + c)] = (double)ws[L.ia + r * 6 + c];
        }
        for (int i = 1; i < n; ++i) {
            const LinkT& l = m.links[i];
            if (l.joint == JF_FIXED) continue;
            int vi = l.v_off;
            SM<R> s = subspace(i);
            R a[36];
            for (int k = 0; k < 36; ++k) a[k] = ws[L.ia + i * 36 + k];
            SF<R> f = m6mul(a, s);
            R diag = sdot(s, f);
            diag = diag + (R)l.armature;
            dw[L.M + vi * nv + vi] = (double)diag;
            int jc = i;
            while (m.links[jc].parent >= 0) {
                int j = m.links[jc].parent;
                f = xtforce(ldx(jc), f);
                jc = j;
                const LinkT& pl = m.links[j];
                int vj = pl.v_off;
                if (pl.joint == JF_HINGE || pl.joint == JF_SLIDE) {
                    double val = (double)sdot(subspace(j), f);
                    dw[L.M + vj * nv + vi] = val;
                    dw[L.M + vi * nv + vj] = val;
                } else if (pl.joint == JF_FREE) {
                    R fc[6] = {f.t.x, f.t.y, f.t.z, f.f.x, f.f.y, f.f.z};
                    for (int kk = 0; kk < 6; ++kk) {
                        dw[L.M + (vj + kk) * nv + vi] = (double)fc[kk];
                        dw[L.M + vi * nv + (vj + kk)] = (double)fc[kk];
                    }
                }
            }
        }
    }

    // Cholesky-Banachiewicz (dynamics.rs:602-622).  false when not PD.
    __device__ bool cholesky(int src, int dst, bool add_damping, double dt) const {
        const int n = m.nv;
        for (int k = 0; k < n * n; ++k) dw[dst + k] = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                double sum = dw[src + i * n + j];
                if (add_damping && i == j) sum = sum + dt * dw[L.dvec + i];
                for (int k = 0; k < j; ++k) sum = sum - dw[dst + i * n + k] * dw[dst + j * n + k];
                if (i == j) {
                    if (!(sum > 0.0)) return false;
                    dw[dst + i * n + i] = __builtin_sqrt(sum);
                } else {
                    dw[dst + i * n + j] = sum / dw[dst + j * n + j];
                }
            }
        return true;
    }
    // cholesky_solve (dynamics.rs:626-648): b in dw[vec], x into dw[res].
    __device__ void chol_solve(int fac, int vec, int res) const {
        const int n = m.nv;
        for (int i = 0; i < n; ++i) {
            double s = dw[vec + i];
            for (int k = 0; k < i; ++k) s = s - dw[fac + i * n + k] * dw[res + k];
            dw[res + i] = s / dw[fac + i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = dw[res + i];
            for (int k = i + 1; k < n; ++k) s = s - dw[fac + k * n + i] * dw[res + k];
            dw[res + i] = s / dw[fac + i * n + i];
        }
    }

    // ---------------- ABA (aba.rs:803-1102, tree.rs:1021-1116) -------------
    // tau: ws[L.tau] generalized applied force; free_applied: whether the
    // free root reads it (disable_penalty_limits).  implicit: joint damping
    // folded into the pivots (Euler).  Result into ws[out].
    __device__ void aba(bool free_applied, bool implicit, R dt, int out) const {
        const int n = m.n_links;
        V3<R> g = mk3((R)m.gravity[0], (R)m.gravity[1], (R)m.gravity[2]);
        for (int i = 0; i < n; ++i) {
            const LinkT& l = m.links[i];
            SM<R> v, c = smz<R>();
            if (l.joint == JF_FREE) {
                int o = l.v_off;
                stx(i, xup(i, false));
                v = {mk3(qd(o), qd(o + 1), qd(o + 2)), mk3(qd(o + 3), qd(o + 4), qd(o + 5))};
            } else if (l.joint == JF_FIXED) {
                Xf<R> x = xup(i, false);
                stx(i, x);
                v = xmotion(x, l.parent >= 0 ? ldm(L.v + l.parent * 6) : smz<R>());
            } else {
                Xf<R> x = xup(i, false);
                stx(i, x);
                SM<R> s = subspace(i);
                stm(L.S + i * 6, s);
                SM<R> sq = s * qd(l.v_off);
                v = xmotion(x, ldm(L.v + l.parent * 6)) + sq;
                c = crm(v, sq);
            }
            stm(L.v + i * 6, v);
            stm(L.c + i * 6, c);
            // body-frame gravity + zero external wrench (tree.rs:1049-1063)
            V3<R> fb = qrot_inv(ori(i), z3<R>() + z3<R>() + g * (R)l.mass);
            V3<R> tb = qrot_inv(ori(i), z3<R>() + z3<R>());
            R a[36];
            for (int k = 0; k < 36; ++k) {
                a[k] = (R)l.si[k];
                ws[L.ia + i * 36 + k] = a[k];
            }
            SF<R> pa = crf(v, m6mul(a, v)) - SF<R>{tb, fb};
            stf(L.pa + i * 6, pa);
        }
        for (int i = n - 1; i >= 1; --i) {
            const LinkT& l = m.links[i];
            int p = l.parent;
            R a[36], out6[36];
            for (int k = 0; k < 36; ++k) a[k] = ws[L.ia + i * 36 + k];
            Xf<R> x = ldx(i);
            SF<R> pa = ldf(L.pa + i * 6);
            SM<R> c = ldm(L.c + i * 6);
            SF<R> paf;
            if (l.joint == JF_FIXED) {
                paf = pa + m6mul(a, c);
            } else {
                SM<R> s = ldm(L.S + i * 6);
                SF<R> ias = m6mul(a, s);
                R dm = implicit ? dt * ((R)l.damping + R(0.0)) : R(0.0);
                R d = (sdot(s, ias) + (R)l.armature) + dm;
                SF<R> ps = pa + m6mul(a, c);
                R tauv = ws[L.tau + l.v_off];
                R fs = (((tauv + R(0.0)) - (R)l.damping * qd(l.v_off)) + R(0.0)) + R(-0.0);
                ws[L.fs + i] = fs;
                R u = fs - sdot(s, ps);
                R qdd = u / d;
                paf = ps + ias * qdd;
                R inv_d = R(1) / d;
                R iv[6] = {ias.t.x, ias.t.y, ias.t.z, ias.f.x, ias.f.y, ias.f.z};
                for (int r = 0; r < 6; ++r)
                    for (int cc = 0; cc < 6; ++cc) a[r * 6 + cc] = a[r * 6 + cc] - (iv[r] * iv[cc]) * inv_d;
                ws[L.d + i] = d;
            }
            m6pullback(a, x, out6);
            for (int k = 0; k < 36; ++k) ws[L.ia + p * 36 + k] = ws[L.ia + p * 36 + k] + out6[k];
            stf(L.pa + p * 6, ldf(L.pa + p * 6) + xtforce(x, paf));
        }
        for (int k = 0; k < m.nv; ++k) ws[out + k] = R(0);
        SM<R> a0 = smz<R>();
        if (m.links[0].joint == JF_FREE) {
            const LinkT& l = m.links[0];
            int o = l.v_off;
            R fr[6];
            for (int k = 0; k < 6; ++k) fr[k] = R(0.0) + (free_applied ? ws[L.tau + o + k] : R(0.0));
            SF<R> pa0 = ldf(L.pa);
            SF<R> rhs = SF<R>{mk3(fr[0], fr[1], fr[2]), mk3(fr[3], fr[4], fr[5])} - pa0;
            R A[6][7];
            for (int r = 0; r < 6; ++r)
                for (int cc = 0; cc < 6; ++cc) A[r][cc] = ws[L.ia + r * 6 + cc];
            R fd = (R)l.free_damping;
            if (fd != R(0)) {
                SM<R> qr{mk3(qd(o), qd(o + 1), qd(o + 2)), mk3(qd(o + 3), qd(o + 4), qd(o + 5))};
                rhs = rhs - SF<R>{qr.w * fd, qr.v * fd};
                R dmr = implicit ? dt * fd : R(0.0);
                for (int k = 0; k < 6; ++k) A[k][k] = A[k][k] + dmr;
            }
            R rv[6] = {rhs.t.x, rhs.t.y, rhs.t.z, rhs.f.x, rhs.f.y, rhs.f.z};
            for (int r = 0; r < 6; ++r) A[r][6] = rv[r];
            // GMat6::solve (aba.rs:614-665): Gauss-Jordan, pivot = last row of
            // maximal |a| under the total order.
            for (int col = 0; col < 6; ++col) {
                int piv = col;
                for (int r = col + 1; r < 6; ++r)
                    if (!(nabs(A[piv][col]) > nabs(A[r][col]))) piv = r;
                if (piv != col)
                    for (int k = 0; k < 7; ++k) { R t = A[col][k]; A[col][k] = A[piv][k]; A[piv][k] = t; }
                R inv = R(1) / A[col][col];
                R prow[7];
                for (int k = 0; k < 7; ++k) prow[k] = A[col][k];
                for (int r = 0; r < 6; ++r) {
                    if (r == col) continue;
                    R factor = A[r][col] * inv;
                    for (int k = col; k < 7; ++k) A[r][k] = A[r][k] - factor * prow[k];
                }
            }
            R x6[6];
            for (int r = 0; r < 6; ++r) x6[r] = A[r][6] / A[r][r];
            a0 = {mk3(x6[0], x6[1], x6[2]), mk3(x6[3], x6[4], x6[5])};
            for (int k = 0; k < 6; ++k) ws[out + o + k] = x6[k];
        }
        stm(L.a, a0);
        for (int i = 1; i < n; ++i) {
            const LinkT& l = m.links[i];
            Xf<R> x = ldx(i);
            SM<R> ap = xmotion(x, ldm(L.a + l.parent * 6));
            if (l.joint == JF_FIXED) {
                stm(L.a + i * 6, ap);
                continue;
            }
            SM<R> prime = ap + ldm(L.c + i * 6);
            R a[36];
            for (int k = 0; k < 36; ++k) a[k] = ws[L.ia + i * 36 + k];
            SF<R> inner = m6mul(a, prime) + ldf(L.pa + i * 6);
            SM<R> s = ldm(L.S + i * 6);
            R qdd = (ws[L.fs + i] - sdot(s, inner)) / ws[L.d + i];
            stm(L.a + i * 6, prime + s * qdd);
            ws[out + l.v_off] = qdd;
        }
    }

    // ---------------- contact rows (solver.rs:2805-3366) -------------------
    // Translational Jacobian of a world point on `link` (jacobian.rs:44-110),
    // into ws[L.jt + 3*slot].
    __device__ void point_jacobian(int link, V3<R> local) const {
        for (int s = 0; s < m.nv; ++s) st3(L.jt + 3 * s, z3<R>());
        V3<R> point = pos(link) + qrot(ori(link), local);
        const LinkT& tl = m.links[link];
        for (int k = 0; k < tl.chain_len; ++k) {
            int i = m.chain[tl.chain_off + k];
            const LinkT& l = m.links[i];
            if (l.joint == JF_FREE) {
                V3<R> r = point - pos(i);
                Qt<R> o = ori(i);
                V3<R> ex = qrot(o, mk3(R(1), R(0), R(0))), ey = qrot(o, mk3(R(0), R(1), R(0))), ez = qrot(o, mk3(R(0), R(0), R(1)));
                st3(L.jt + 3 * (l.v_off + 0), cross3(ex, r));
                st3(L.jt + 3 * (l.v_off + 1), cross3(ey, r));
                st3(L.jt + 3 * (l.v_off + 2), cross3(ez, r));
                st3(L.jt + 3 * (l.v_off + 3), ex);
                st3(L.jt + 3 * (l.v_off + 4), ey);
                st3(L.jt + 3 * (l.v_off + 5), ez);
            } else if (l.joint == JF_HINGE) {
                Qt<R> po = ori(l.parent);
                V3<R> jw = pos(l.parent) + qrot(po, mk3((R)l.pj[0], (R)l.pj[1], (R)l.pj[2]));
                V3<R> aw = qrot(po, mk3((R)l.axis[0], (R)l.axis[1], (R)l.axis[2]));
                st3(L.jt + 3 * l.v_off, cross3(aw, point - jw));
            } else if (l.joint == JF_SLIDE) {
                st3(L.jt + 3 * l.v_off, qrot(ori(l.parent), mk3((R)l.axis[0], (R)l.axis[1], (R)l.axis[2])));
            }
        }
    }

    // tree_body_invweight0 (solver.rs:3719-3754), translational part, cached
    // per link for the step (it depends on the link only).
    __device__ R invweight(int link) const {
        if (iw[L.iwf + link]) return ws[L.iw + link];
        point_jacobian(link, z3<R>());
        const int nv = m.nv;
        R tr = R(0.0);
        for (int axis = 0; axis < 3; ++axis) {
            for (int s = 0; s < nv; ++s) dw[L.dvec + s] = (double)ws[L.jt + 3 * s + axis];
            chol_solve(L.Ld, L.dvec, L.dres);
            Sum<R> sm;
            for (int s = 0; s < nv; ++s) sm.add(ws[L.jt + 3 * s + axis] * (R)dw[L.dres + s]);
            tr = tr + sm.acc;
        }
        R v = tr / R(3.0);
        ws[L.iw + link] = v;
        iw[L.iwf + link] = 1;
        return v;
    }

    // Builds rows, Delassus, PGS; leaves the generalized contact force in
    // ws[L.delta].  Returns false if a mass matrix is not positive definite.
    __device__ bool solve_contacts(int nc, R dt, int* nrows_out) const {
        const int nv = m.nv;
        for (int s = 0; s < nv; ++s) ws[L.delta + s] = R(0.0);
        *nrows_out = 0;
        // Active contacts (penetration beyond the force-free gap).
        bool any = false;
        for (int k = 0; k < nc; ++k)
            if (ws[L.cgeo + 9 * k + 6] - ws[L.cgeo + 9 * k + 8] > R(0.0)) any = true;
        if (!any) return true;
        // Response factor (M + dt*B for Euler) and the diagApprox factor (M).
        mass_matrix();
        bool implicit = m.integrator != 2;
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            if (l.joint == JF_FREE) for (int k = 0; k < 6; ++k) dw[L.dvec + l.v_off + k] = l.free_damping;
            else if (l.joint == JF_HINGE || l.joint == JF_SLIDE) dw[L.dvec + l.v_off] = l.damping + 0.0;
        }
        if (!cholesky(L.M, L.Li, implicit && (double)dt != 0.0, (double)dt)) return false;
        if (!cholesky(L.M, L.Ld, false, 0.0)) return false;
        // Free velocity: ABA at the current state, user forces only.
        for (int s = 0; s < nv; ++s) ws[L.tau + s] = (R)qfrc[s];
        aba(true, implicit, dt, L.qacc_free);
        for (int i = 0; i < m.n_links; ++i) iw[L.iwf + i] = 0;
        int nr = 0;
        for (int k = 0; k < nc; ++k) {
            int o = L.cgeo + 9 * k;
            R pen = ws[o + 6] - ws[o + 8];
            if (pen <= R(0.0)) continue;
            const GeomT& ga = m.geoms[iw[L.cids + 3 * k]];
            const GeomT& gb = m.geoms[iw[L.cids + 3 * k + 1]];
            V3<R> cp = ld3(o), n = ld3(o + 3);
            R mu = (R)ga.friction < (R)gb.friction ? (R)ga.friction : (R)gb.friction;
            // tangent_basis (world.rs:3104-3111)
            V3<R> ref = nabs(n.z) < R(0.9) ? mk3(R(0), R(0), R(1)) : mk3(R(1), R(0), R(0));
            V3<R> t1 = unit3(cross3(ref, n));
            V3<R> t2 = cross3(n, t1);
            V3<R> dirs[4] = {n + t1 * mu, n - t1 * mu, n + t2 * mu, n - t2 * mu};
            for (int r = 0; r < 4; ++r) iw[L.rinfo + 3 * (nr + r)] = 0;
            for (int side = 0; side < 2; ++side) {
                const GeomT& g = side == 0 ? ga : gb;
                if (g.link < 0) continue;
                R sign = side == 0 ? R(1.0) : R(-1.0);
                V3<R> local = qrot_inv(ori(g.link), cp - pos(g.link));
                point_jacobian(g.link, local);
                for (int r = 0; r < 4; ++r) {
                    int row = nr + r;
                    int cidx = iw[L.rinfo + 3 * row];
                    iw[L.rinfo + 3 * row + 1 + cidx] = g.link;
                    iw[L.rinfo + 3 * row] = cidx + 1;
                    int co = L.coef + (row * 2 + cidx) * nv;
                    for (int s = 0; s < nv; ++s) ws[co + s] = dot3(ld3(L.jt + 3 * s), dirs[r]) * sign;
                }
            }
            // Per-row constants (solver.rs:3067-3120).
            double sr0 = ga.solref[0] <= gb.solref[0] ? ga.solref[0] : gb.solref[0];
            double sr1 = ga.solref[1] <= gb.solref[1] ? ga.solref[1] : gb.solref[1];
            if (ga.solref[0] < 0.0 || gb.solref[0] < 0.0) {
                if (!(ga.solref[0] < 0.0 && gb.solref[0] < 0.0)) {
                    const GeomT& gd = ga.solref[0] < 0.0 ? ga : gb;
                    sr0 = gd.solref[0];
                    sr1 = gd.solref[1];
                }
            }
            R dmin = (R)(ga.solimp[0] < gb.solimp[0] ? ga.solimp[0] : gb.solimp[0]);
            R dmax = (R)(ga.solimp[1] < gb.solimp[1] ? ga.solimp[1] : gb.solimp[1]);
            R width = (R)(ga.solimp[2] < gb.solimp[2] ? ga.solimp[2] : gb.solimp[2]);
            R mid = (R)ga.solimp[3];
            int power = ga.power;
            // effective_solimp (solver.rs:319-328)
            R lo = R(1.0e-4), hi = R(0.9999);
            auto clampi = [&](R v) { if (v < lo) v = lo; if (v > hi) v = hi; return v; };
            dmin = clampi(dmin); dmax = clampi(dmax); mid = clampi(mid);
            width = width > R(0.0) ? width : R(0.0);
            if (power < 1) power = 1;
            auto impedance = [&](R position) {
                R x = (position - R(0.0)) / width;
                if (x < R(0.0)) x = -x;
                R y;
                if (x <= R(0.0)) y = R(0.0);
                else if (x >= R(1.0)) y = R(1.0);
                else if (x < mid) {
                    R b = x / mid, pw = R(1.0);
                    for (int k2 = 0; k2 < power; ++k2) pw = pw * b;
                    y = mid * pw;
                } else {
                    R b = (R(1.0) - x) / (R(1.0) - mid), pw = R(1.0);
                    for (int k2 = 0; k2 < power; ++k2) pw = pw * b;
                    y = R(1.0) - (R(1.0) - mid) * pw;
                }
                return dmin + (dmax - dmin) * y;
            };
            R imp = impedance(-pen);
            // safe_solref (solver.rs:331-337) and reference_accel (353-384)
            R tc = (R)sr0, dr = (R)sr1;
            bool direct = tc < R(0.0);
            if (!direct && tc < R(2.0) * dt) tc = R(2.0) * dt;
            R position = -pen;
            R dref = impedance(position);
            R bcoef, kcoef;
            if (direct) {
                bcoef = -dr / dmax;
                kcoef = -tc / (dmax * dmax);
            } else {
                bcoef = R(2.0) / (dmax * tc);
                kcoef = dref / (dmax * dmax * tc * tc * dr * dr);
            }
            for (int r = 0; r < 4; ++r) {
                int row = nr + r;
                int ncomp = iw[L.rinfo + 3 * row];
                R tr = R(0.0);
                Sum<R> vel, fvel;
                for (int cidx = 0; cidx < ncomp; ++cidx) {
                    int co = L.coef + (row * 2 + cidx) * nv;
                    tr = tr + invweight(iw[L.rinfo + 3 * row + 1 + cidx]);
                    Sum<R> sv, sf;
                    for (int s = 0; s < nv; ++s) sv.add(ws[co + s] * qd(s));
                    for (int s = 0; s < nv; ++s) sf.add(ws[co + s] * (ws[L.qacc_free + s] * dt));
                    vel.add(R(0.0) + sv.acc);
                    fvel.add(R(0.0) + sf.acc);
                }
                R diag_approx = tr * (R(1.0) + mu * mu);
                R base = (R(1.0) - imp) / imp * diag_approx;
                ws[L.rreg + row] = R(2.0) * mu * mu * base;
                R refacc = -bcoef * vel.acc - kcoef * position;
                ws[L.rbias + row] = fvel.acc - refacc * dt;
            }
            nr += 4;
        }
        *nrows_out = nr;
        if (nr == 0) return true;
        // Delassus A = J (M+dtB)^-1 Jᵀ by columns (solver.rs:3045-3065).
        const int cap = m.cap_r;
        for (int col = 0; col < nr; ++col) {
            int ncomp = iw[L.rinfo + 3 * col];
            for (int s = 0; s < nv; ++s) ws[L.gres + s] = R(0.0);
            for (int cidx = 0; cidx < ncomp; ++cidx) {
                int co = L.coef + (col * 2 + cidx) * nv;
                for (int s = 0; s < nv; ++s) dw[L.dvec + s] = (double)(ws[co + s] * R(1.0));
                chol_solve(L.Li, L.dvec, L.dres);
                for (int s = 0; s < nv; ++s) ws[L.gres + s] = ws[L.gres + s] + (R)dw[L.dres + s];
            }
            for (int row = 0; row < nr; ++row) {
                int nrc = iw[L.rinfo + 3 * row];
                Sum<R> outer;
                for (int cidx = 0; cidx < nrc; ++cidx) {
                    int co = L.coef + (row * 2 + cidx) * nv;
                    Sum<R> sm;
                    for (int s = 0; s < nv; ++s) sm.add(ws[co + s] * ws[L.gres + s]);
                    outer.add(R(0.0) + sm.acc);
                }
                ws[L.A + row * cap + col] = outer.acc;
            }
        }
        for (int row = 0; row < nr; ++row) {
            ws[L.rdiag + row] = ws[L.A + row * cap + row] + ws[L.rreg + row];
            ws[L.f + row] = R(0.0);
        }
        // Projected Gauss-Seidel (solver.rs:3269-3331, 3838-3861).
        for (int it = 0; it < m.iterations; ++it)
            for (int i = 0; i < nr; ++i) {
                Sum<R> sm;
                for (int col = 0; col < nr; ++col) sm.add(ws[L.A + i * cap + col] * ws[L.f + col]);
                R fi = ws[L.f + i];
                R residual = (ws[L.rbias + i] + sm.acc) + ws[L.rreg + i] * fi;
                ws[L.f + i] = nmax(fi - residual / ws[L.rdiag + i], R(0.0));
            }
        // Generalized force (solver.rs:3345-3364).
        for (int i = 0; i < nr; ++i) {
            R force = ws[L.f + i] / dt;
            int ncomp = iw[L.rinfo + 3 * i];
            for (int cidx = 0; cidx < ncomp; ++cidx) {
                int co = L.coef + (i * 2 + cidx) * nv;
                for (int s = 0; s < nv; ++s) ws[L.delta + s] = ws[L.delta + s] + ws[co + s] * force;
            }
        }
        return true;
    }

    // ---------------- one World::step --------------------------------------
    // Precondition: poses and link velocities describe the current state.
    __device__ int step() const {
        const R dt = (R)m.dt;
        const int nv = m.nv;
        int np = active_pairs();
        int nc = 0;
        bool constraint = m.mode == 1;
        for (int s = 0; s < nv; ++s) ws[L.delta + s] = R(0.0);
        if (constraint) {
            nc = narrowphase(np);
            int nr;
            if (!solve_contacts(nc, dt, &nr)) return 3;  // NEWT_ERR_NOT_PD
        }
        // step_trees_euler (world.rs:1952-2051): applied + (0 + delta).
        for (int s = 0; s < nv; ++s) ws[L.tau + s] = (R)qfrc[s] + (R(0.0) + ws[L.delta + s]);
        aba(constraint, true, dt, L.delta);
        for (int s = 0; s < nv; ++s) qacc[s] = (double)ws[L.delta + s];
        // euler_step_with_workspace (tree.rs:1407-1418) + integrate_tree_positions
        // (1422-1475).  Velocity and position accumulation is FP64 in both
        // steps.
        const double dtd = m.dt;
        for (int s = 0; s < nv; ++s) qvel[s] = qvel[s] + (double)ws[L.delta + s] * dtd;
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            if (l.joint == JF_HINGE || l.joint == JF_SLIDE) {
                qpos[l.q_off] = qpos[l.q_off] + qvel[l.v_off] * dtd;
            } else if (l.joint == JF_FREE) {
                int qo = l.q_off, vo = l.v_off;
                Qt<double> o{qpos[qo + 3], qpos[qo + 4], qpos[qo + 5], qpos[qo + 6]};
                V3<double> w = mk3(qvel[vo], qvel[vo + 1], qvel[vo + 2]);
                V3<double> vel = mk3(qvel[vo + 3], qvel[vo + 4], qvel[vo + 5]);
                V3<double> dp = qrot(o, vel) * dtd;
                qpos[qo] = qpos[qo] + dp.x;
                qpos[qo + 1] = qpos[qo + 1] + dp.y;
                qpos[qo + 2] = qpos[qo + 2] + dp.z;
                Qt<double> nx = qintegrate(o, w, dtd);
                qpos[qo + 3] = nx.x; qpos[qo + 4] = nx.y;
                qpos[qo + 5] = nx.z; qpos[qo + 6] = nx.w;
            }
        }
        // refresh_broadphase_for_query on the post-step state (world.rs:1523)
        forward_kinematics();
        link_velocities();
        populate(dt, false);
        iw[L.iwf + m.n_links] = np;  // pair count of this step
        iw[L.iwf + m.n_links + 1] = nc;
        return 0;
    }
};

}  // namespace newt
