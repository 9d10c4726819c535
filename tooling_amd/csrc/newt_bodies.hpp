// newt_bodies.hpp — the free-body step, one lane per world.
//
// A model of free bodies (World::add_body, newt/src/world.rs:950) steps as
// World::step does for its bodies (world.rs:1459-1537): broadphase pairs and
// narrowphase contacts (shared with the tree step: World<double>::populate,
// active_pairs, narrowphase on the bodies' pose slots), the free-body
// constraint solve solve_free_bodies (PGS, pyramidal cone, condim 3;
// newt/src/solver.rs:703-1257) into one world wrench per body, then
// step_bodies_euler (world.rs:1790-1817) or step_bodies (RK4 with the
// wrenches held over the step, 1705-1785), and the broadphase refresh.
//
// Arithmetic is FP64 with the reference's operation order (the unit is
// compiled without FP contraction), so it reproduces the FP64 oracle bit for
// bit.  Per-world arrays are lane-interleaved like the tree step's.
#pragma once

#include "../../include/newt_gpu.h"
#include "newt_step.hpp"

namespace newt {

struct FreeBodies {
    const World<double>& w;

    NEWT_HD const BodyT& body(int b) const { return w.m.bodies[b]; }
    NEWT_HD V3<double> bpos(int b) const {
        const int o = body(b).q_off;
        return mk3(w.qpos[o], w.qpos[o + 1], w.qpos[o + 2]);
    }
    NEWT_HD Qt<double> bori(int b) const {
        const int o = body(b).q_off;
        return Qt<double>{w.qpos[o + 3], w.qpos[o + 4], w.qpos[o + 5], w.qpos[o + 6]};
    }
    NEWT_HD V3<double> bvel(int b) const {
        const int o = body(b).v_off;
        return mk3(w.qvel[o], w.qvel[o + 1], w.qvel[o + 2]);
    }
    NEWT_HD V3<double> bomega(int b) const {
        const int o = body(b).v_off;
        return mk3(w.qvel[o + 3], w.qvel[o + 4], w.qvel[o + 5]);
    }
    NEWT_HD M33<double> ldm(int off) const {
        M33<double> a;
        for (int k = 0; k < 9; ++k) a.m[k] = w.dw[off + k];
        return a;
    }
    static NEWT_HD M33<double> tbl(const double* p) {
        M33<double> a;
        for (int k = 0; k < 9; ++k) a.m[k] = p[k];
        return a;
    }
    // Mat3 product and transpose (math.rs:523-537), sums from 0.0.
    static NEWT_HD M33<double> mm(const M33<double>& a, const M33<double>& b) {
        M33<double> o;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += a.m[r * 3 + k] * b.m[k * 3 + c];
                o.m[r * 3 + c] = s;
            }
        return o;
    }
    static NEWT_HD M33<double> tr(const M33<double>& a) {
        M33<double> o;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) o.m[r * 3 + c] = a.m[c * 3 + r];
        return o;
    }
    NEWT_HD V3<double> gravity_of(int b) const {
        const V3<double> g = mk3(w.m.gravity[0], w.m.gravity[1], w.m.gravity[2]);
        const double s = body(b).gravity_scale;
        return s == 1.0 ? g : g * s;
    }

    // Pose slot n_links + b = body b's pose (geom_world_pose reads it).
    NEWT_HD void set_poses() const {
        for (int b = 0; b < w.m.n_bodies; ++b) w.set_pose(w.m.n_links + b, bpos(b), bori(b));
    }

    // ---------------- rows (solver.rs:1892-2069, 2131-2154, 4012-4036) -----
    // Row ri: dir at brow + 7 ri, then reg, diag, bias, impulse.  Block k
    // (rows 4k..4k+3): arm_a, arm_b, h01, h23 at bblk + 8k; bodies at bbi + 2k.
    NEWT_HD int ba(int k) const { return w.iw[w.L.bbi + 2 * k]; }
    NEWT_HD int bb(int k) const { return w.iw[w.L.bbi + 2 * k + 1]; }
    NEWT_HD V3<double> rdir(int ri) const { return w.kld3(w.L.brow + 7 * ri); }
    NEWT_HD double& rreg(int ri) const { return w.dw[w.L.brow + 7 * ri + 3]; }
    NEWT_HD double& rdiag(int ri) const { return w.dw[w.L.brow + 7 * ri + 4]; }
    NEWT_HD double& rbias(int ri) const { return w.dw[w.L.brow + 7 * ri + 5]; }
    NEWT_HD double& rimp(int ri) const { return w.dw[w.L.brow + 7 * ri + 6]; }
    NEWT_HD V3<double> arm_a(int k) const { return w.kld3(w.L.bblk + 8 * k); }
    NEWT_HD V3<double> arm_b(int k) const { return w.kld3(w.L.bblk + 8 * k + 3); }

    // Impulse accumulator of body i: dv_lin_world, dw_ang_body.
    NEWT_HD V3<double> bdv(int i) const { return w.kld3(w.L.bbd + 6 * i); }
    NEWT_HD V3<double> bdw(int i) const { return w.kld3(w.L.bbd + 6 * i + 3); }

    // row_residual with explicit accumulators of the row's two bodies.
    NEWT_HD double residual_of(V3<double> dir, int a, int b, V3<double> arma, V3<double> armb, V3<double> dva,
                               V3<double> dwa, V3<double> dvb, V3<double> dwb) const {
        double s = 0.0;
        if (a >= 0) {
            V3<double> wa = qrot(bori(a), dwa);
            V3<double> vp = dva + cross3(wa, arma);
            s += dot3(vp, dir);
        }
        if (b >= 0) {
            V3<double> wb = qrot(bori(b), dwb);
            V3<double> vp = dvb + cross3(wb, armb);
            s -= dot3(vp, dir);
        }
        return s;
    }
    NEWT_HD double residual(int ri) const {
        const int k = ri >> 2;
        const int a = ba(k), b = bb(k);
        V3<double> z = z3<double>();
        return residual_of(rdir(ri), a, b, arm_a(k), arm_b(k), a >= 0 ? bdv(a) : z, a >= 0 ? bdw(a) : z,
                           b >= 0 ? bdv(b) : z, b >= 0 ? bdw(b) : z);
    }
    // apply_impulse_delta of one side into (dv, dw).
    NEWT_HD void side_delta(int i, V3<double> dir, V3<double> arm, double df, V3<double>* dv, V3<double>* dwv) const {
        V3<double> f = dir * df;
        V3<double> tq = cross3(arm, f);
        *dv = *dv + f / body(i).mass;
        *dwv = *dwv + qrot_inv(bori(i), mv(ldm(w.L.biw + 9 * i), tq));
    }
    NEWT_HD void apply(int ri, double df) const {
        if (df == 0.0) return;
        const int k = ri >> 2;
        const int a = ba(k), b = bb(k);
        const V3<double> dir = rdir(ri);
        if (a >= 0) {
            V3<double> dv = bdv(a), dwv = bdw(a);
            side_delta(a, dir, arm_a(k), df, &dv, &dwv);
            w.kst3(w.L.bbd + 6 * a, dv);
            w.kst3(w.L.bbd + 6 * a + 3, dwv);
        }
        if (b >= 0) {
            V3<double> dv = bdv(b), dwv = bdw(b);
            side_delta(b, -dir, arm_b(k), df, &dv, &dwv);
            w.kst3(w.L.bbd + 6 * b, dv);
            w.kst3(w.L.bbd + 6 * b + 3, dwv);
        }
    }
    // pgs_step_non_negative (solver.rs:1448-1466).
    NEWT_HD void step_non_negative(int ri) const {
        const double f = rimp(ri);
        const double res = residual(ri) + rreg(ri) * f + rbias(ri);
        double delta = -res / rdiag(ri);
        const double nf = f + delta;
        const double pr = nf < 0.0 ? 0.0 : nf;
        delta = pr - f;
        rimp(ri) = pr;
        apply(ri, delta);
    }
    // pgs_step_pyramidal_pair (solver.rs:1476-1520).
    NEWT_HD void step_pair(int ri, double h01) const {
        const int rj = ri + 1;
        const double fi = rimp(ri), fj = rimp(rj);
        const double res_i = residual(ri) + rreg(ri) * fi + rbias(ri);
        const double res_j = residual(rj) + rreg(rj) * fj + rbias(rj);
        const double h00 = rdiag(ri), h11 = rdiag(rj);
        const double det = h00 * h11 - h01 * h01;
        if (det <= 0.0) {
            step_non_negative(ri);
            step_non_negative(rj);
            return;
        }
        const double di = (-res_i * h11 + h01 * res_j) / det;
        const double dj = (h01 * res_i - h00 * res_j) / det;
        const double ni = fi + di, nj = fj + dj;
        double a, b;
        if (ni >= 0.0 && nj >= 0.0) {
            a = ni;
            b = nj;
        } else if (ni < 0.0 && nj < 0.0) {
            a = 0.0;
            b = 0.0;
        } else if (ni < 0.0) {
            a = 0.0;
            b = nmax(fj - res_j / h11, 0.0);
        } else {
            a = nmax(fi - res_i / h00, 0.0);
            b = 0.0;
        }
        rimp(ri) = a;
        rimp(rj) = b;
        apply(ri, a - fi);
        apply(rj, b - fj);
    }

    // ---------------- solve_free_bodies (solver.rs:703-1257) ---------------
    // Per-body world wrenches (force, torque) into bwr.
    NEWT_HD void solve(int nc) const {
        const int nb = w.m.n_bodies;
        const double dt = w.m.dt;
        for (int i = 0; i < 6 * nb; ++i) w.dw[w.L.bwr + i] = 0.0;
        if (nc == 0 || dt <= 0.0) return;
        // I_world^-1 = R I^-1 R^T per body.
        for (int i = 0; i < nb; ++i) {
            M33<double> r = qmat(bori(i));
            M33<double> iw = mm(mm(r, tbl(body(i).inv_inertia)), tr(r));
            for (int k = 0; k < 9; ++k) w.dw[w.L.biw + 9 * i + k] = iw.m[k];
        }
        int nblk = 0;
        for (int c = 0; c < nc; ++c) {
            const int o = w.L.cgeo + 9 * c;
            const GeomT& ga = w.m.geoms[w.iw[w.L.cids + 3 * c]];
            const GeomT& gb = w.m.geoms[w.iw[w.L.cids + 3 * c + 1]];
            const int a = ga.body, b = gb.body;
            if (a < 0 && b < 0) continue;
            const double pen = w.dw[o + 6] - w.dw[o + 8];
            if (pen <= 0.0) continue;
            const double mu = ga.friction < gb.friction ? ga.friction : gb.friction;
            const V3<double> cp = w.kld3(o), n = w.kld3(o + 3);
            // contact_friction_axes without anisotropy: tangent_basis (world.rs:3104-3111)
            const V3<double> ref = nabs(n.z) < 0.9 ? mk3(0.0, 0.0, 1.0) : mk3(1.0, 0.0, 0.0);
            const V3<double> t1 = unit3(cross3(ref, n));
            const V3<double> t2 = cross3(n, t1);
            const V3<double> arma = a >= 0 ? cp - bpos(a) : z3<double>();
            const V3<double> armb = b >= 0 ? cp - bpos(b) : z3<double>();
            const int k = nblk++;
            w.iw[w.L.bbi + 2 * k] = a;
            w.iw[w.L.bbi + 2 * k + 1] = b;
            w.kst3(w.L.bblk + 8 * k, arma);
            w.kst3(w.L.bblk + 8 * k + 3, armb);
            // contact_row_directions (solver.rs:3401-3417)
            const V3<double> dirs[4] = {n + t1 * mu, n - t1 * mu, n + t2 * mu, n - t2 * mu};
            // combine_solref (geom.rs:156-185), combine_solimp (solver.rs:242-250)
            double sr0 = ga.solref[0] <= gb.solref[0] ? ga.solref[0] : gb.solref[0];
            double sr1 = ga.solref[1] <= gb.solref[1] ? ga.solref[1] : gb.solref[1];
            if (ga.solref[0] < 0.0 || gb.solref[0] < 0.0) {
                if (!(ga.solref[0] < 0.0 && gb.solref[0] < 0.0)) {
                    const GeomT& gd = ga.solref[0] < 0.0 ? ga : gb;
                    sr0 = gd.solref[0];
                    sr1 = gd.solref[1];
                }
            }
            const double si[4] = {ga.solimp[0] < gb.solimp[0] ? ga.solimp[0] : gb.solimp[0],
                                  ga.solimp[1] < gb.solimp[1] ? ga.solimp[1] : gb.solimp[1],
                                  ga.solimp[2] < gb.solimp[2] ? ga.solimp[2] : gb.solimp[2], ga.solimp[3]};
            const SolImpE se = solimp_effective(si, ga.power);
            const double imp = impedance_at(se, -pen);
            for (int r = 0; r < 4; ++r) {
                const int ri = 4 * k + r;
                const V3<double> dir = dirs[r];
                w.kst3(w.L.brow + 7 * ri, dir);
                // row_body_diagonal, linear row
                double aii = 0.0;
                if (a >= 0) {
                    V3<double> rc = cross3(arma, dir);
                    aii += dot3(rc, mv(ldm(w.L.biw + 9 * a), rc)) + 1.0 / body(a).mass;
                }
                if (b >= 0) {
                    V3<double> rc = cross3(armb, dir);
                    aii += dot3(rc, mv(ldm(w.L.biw + 9 * b), rc)) + 1.0 / body(b).mass;
                }
                // contact_regularization (solver.rs:1930-1974), pyramidal condim 3
                double reg = 0.0;
                if (imp > 0.0) {
                    double tran = 0.0;
                    if (a >= 0) tran += 1.0 / body(a).mass;
                    if (b >= 0) tran += 1.0 / body(b).mass;
                    const double base = (1.0 - imp) / imp * (tran * (1.0 + mu * mu));
                    reg = 2.0 * mu * mu * base;
                }
                rreg(ri) = reg;
                rdiag(ri) = aii + reg;
                // normal_velocity_at_point and the gravity-only free step
                V3<double> va = z3<double>(), vb = z3<double>();
                if (a >= 0) va = bvel(a) + cross3(qrot(bori(a), bomega(a)), arma);
                if (b >= 0) vb = bvel(b) + cross3(qrot(bori(b), bomega(b)), armb);
                const double vcur = dot3(va - vb, dir);
                double dvf = 0.0;
                if (a >= 0) {
                    V3<double> dwz = qrot(bori(a), z3<double>());
                    dvf += dot3(gravity_of(a) * dt + cross3(dwz, arma), dir);
                }
                if (b >= 0) {
                    V3<double> dwz = qrot(bori(b), z3<double>());
                    dvf -= dot3(gravity_of(b) * dt + cross3(dwz, armb), dir);
                }
                const double aref = reference_accel(-pen, vcur, sr0, sr1, se, dt);
                rbias(ri) = dvf - aref * dt;
                rimp(ri) = 0.0;
            }
            // row_cross_response_into (solver.rs:1604-1618): a unit impulse on
            // row 4k+1 (4k+3) seen by row 4k (4k+2).
            for (int p = 0; p < 2; ++p) {
                const V3<double> applied = dirs[2 * p + 1];
                V3<double> dva = z3<double>(), dwa = z3<double>(), dvb = z3<double>(), dwb = z3<double>();
                if (a >= 0) side_delta(a, applied, arma, 1.0, &dva, &dwa);
                if (b >= 0) side_delta(b, -applied, armb, 1.0, &dvb, &dwb);
                w.dw[w.L.bblk + 8 * k + 6 + p] = residual_of(dirs[2 * p], a, b, arma, armb, dva, dwa, dvb, dwb);
            }
        }
        if (nblk == 0) return;
        for (int i = 0; i < 6 * nb; ++i) w.dw[w.L.bbd + i] = 0.0;
        for (int it = 0; it < w.m.iterations; ++it)
            for (int k = 0; k < nblk; ++k) {
                step_non_negative(4 * k);
                step_pair(4 * k, w.dw[w.L.bblk + 8 * k + 6]);
                step_pair(4 * k + 2, w.dw[w.L.bblk + 8 * k + 7]);
            }
        // Per-body wrenches, rows in order (solver.rs:1198-1235).
        for (int ri = 0; ri < 4 * nblk; ++ri) {
            const double f = rimp(ri);
            if (f == 0.0) continue;
            const int k = ri >> 2;
            const int a = ba(k), b = bb(k);
            const V3<double> fa = rdir(ri) * (f / dt);
            if (a >= 0) {
                const int o = w.L.bwr + 6 * a;
                w.kst3(o, w.kld3(o) + fa);
                w.kst3(o + 3, w.kld3(o + 3) + cross3(arm_a(k), fa));
            }
            if (b >= 0) {
                const V3<double> fb = -fa;
                const int o = w.L.bwr + 6 * b;
                w.kst3(o, w.kld3(o) + fb);
                w.kst3(o + 3, w.kld3(o + 3) + cross3(arm_b(k), fb));
            }
        }
    }

    // ---------------- integration (world.rs:1705-1817, 3223-3286) ---------
    struct St {
        V3<double> p, v;
        Qt<double> q;
        V3<double> om;
    };
    struct Dv {
        V3<double> dp, dv;
        Qt<double> dq;
        V3<double> dw;
    };
    // evaluate (world.rs:3223-3245)
    NEWT_HD Dv evaluate(int b, const St& s, V3<double> f, V3<double> tq) const {
        const BodyT& bt = body(b);
        Dv d;
        d.dv = gravity_of(b) + f / bt.mass;
        const V3<double> iwv = mv(tbl(bt.inertia), s.om);
        const V3<double> gyro = -cross3(s.om, iwv);
        const V3<double> tb = qrot_inv(s.q, tq);
        d.dw = mv(tbl(bt.inv_inertia), tb + gyro);
        d.dq = qderiv(s.q, s.om);
        d.dp = s.v;
        return d;
    }
    static NEWT_HD Qt<double> qadd(Qt<double> a, Qt<double> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
    static NEWT_HD Qt<double> qsc(Qt<double> a, double s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
    static NEWT_HD St advance(const St& s, const Dv& d, double h) {
        return St{s.p + d.dp * h, s.v + d.dv * h, qadd(s.q, qsc(d.dq, h)), s.om + d.dw * h};
    }
    // clamp_velocity (world.rs:3273-3286)
    static NEWT_HD V3<double> clampv(V3<double> v, int has, double cap) {
        if (!has) return v;
        if (cap <= 0.0) return z3<double>();
        const double speed = len3(v);
        return speed > cap ? unit3(v) * cap : v;
    }

    NEWT_HD void integrate() const {
        const double dt = w.m.dt;
        for (int b = 0; b < w.m.n_bodies; ++b) {
            const BodyT& bt = body(b);
            const St s0{bpos(b), bvel(b), bori(b), bomega(b)};
            const V3<double> f = w.kld3(w.L.bwr + 6 * b), tq = w.kld3(w.L.bwr + 6 * b + 3);
            const Dv k1 = evaluate(b, s0, f, tq);
            St n;
            if (w.m.integrator == NEWT_INTEGRATOR_RK4) {
                const Dv k2 = evaluate(b, advance(s0, k1, dt * 0.5), f, tq);
                const Dv k3 = evaluate(b, advance(s0, k2, dt * 0.5), f, tq);
                const Dv k4 = evaluate(b, advance(s0, k3, dt), f, tq);
                const double sixth = 1.0 / 6.0;
                const V3<double> dp = (((k1.dp + k2.dp * 2.0) + k3.dp * 2.0) + k4.dp) * sixth;
                const V3<double> dv = (((k1.dv + k2.dv * 2.0) + k3.dv * 2.0) + k4.dv) * sixth;
                const Qt<double> dq = qsc(qadd(qadd(qadd(k1.dq, qsc(k2.dq, 2.0)), qsc(k3.dq, 2.0)), k4.dq), sixth);
                const V3<double> dwv = (((k1.dw + k2.dw * 2.0) + k3.dw * 2.0) + k4.dw) * sixth;
                n.p = s0.p + dp * dt;
                n.v = clampv(s0.v + dv * dt, bt.has_max_lin, bt.max_lin);
                n.q = qnormalize(qadd(s0.q, qsc(dq, dt)));
                n.om = clampv(s0.om + dwv * dt, bt.has_max_ang, bt.max_ang);
            } else {
                const V3<double> v = s0.v + k1.dv * dt;
                const V3<double> om = s0.om + k1.dw * dt;
                n.p = s0.p + v * dt;
                n.q = qintegrate(s0.q, om, dt);
                n.v = clampv(v, bt.has_max_lin, bt.max_lin);
                n.om = clampv(om, bt.has_max_ang, bt.max_ang);
            }
            const int qo = bt.q_off, vo = bt.v_off;
            w.qpos[qo] = n.p.x; w.qpos[qo + 1] = n.p.y; w.qpos[qo + 2] = n.p.z;
            w.qpos[qo + 3] = n.q.x; w.qpos[qo + 4] = n.q.y; w.qpos[qo + 5] = n.q.z; w.qpos[qo + 6] = n.q.w;
            w.qvel[vo] = n.v.x; w.qvel[vo + 1] = n.v.y; w.qvel[vo + 2] = n.v.z;
            w.qvel[vo + 3] = n.om.x; w.qvel[vo + 4] = n.om.y; w.qvel[vo + 5] = n.om.z;
            w.qacc[vo] = k1.dv.x; w.qacc[vo + 1] = k1.dv.y; w.qacc[vo + 2] = k1.dv.z;
            w.qacc[vo + 3] = k1.dw.x; w.qacc[vo + 4] = k1.dw.y; w.qacc[vo + 5] = k1.dw.z;
        }
    }

    // World::step for a free-body model; 0 or a newt_status fault code.
    NEWT_HD int step() const {
        const int np = w.active_pairs();
        int nc = 0;
        if (w.m.mode == NEWT_SOLVER_PGS) {
            nc = w.narrowphase(np);
            if (nc > w.m.cap_c) return 8;  // NEWT_ERR_CAPACITY
        }
        w.iw[w.L.iwf] = np;
        w.iw[w.L.iwf + 1] = nc;
        solve(nc);
        integrate();
        set_poses();
        w.populate(w.m.dt, false);
        return 0;
    }
};

}  // namespace newt
