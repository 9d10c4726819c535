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
    static NEWT_HD M33<double> tbl(const double* p) { return m33_load(p); }
    static NEWT_HD M33<double> mm(const M33<double>& a, const M33<double>& b) { return m33_mul(a, b); }
    static NEWT_HD M33<double> tr(const M33<double>& a) { return m33_tr(a); }
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
    // solve_free_body_newton_impulses (solver.rs:1295-1351) on a model with
    // trees (free-body models run it on the wavefront step): A[i][j] =
    // row_residual of row i after apply_impulse_delta of a unit impulse on
    // row j, into the tree solve's Delassus storage (done with by now), H =
    // A + diag(reg), lin = the row biases, and the World's nonnegative
    // Newton (newton.rs:275-556) into the rows' impulses.  0, or
    // NEWT_ERR_NEWTON_FAIL when H is not positive definite.
    NEWT_HD int newton_impulses(int nr) const {
        const int cap = w.m.cap_r;
        const V3<double> z = z3<double>();
        for (int j = 0; j < nr; ++j) {
            const int kj = j >> 2, ja = ba(kj), jb = bb(kj);
            V3<double> dva = z, dwa = z, dvb = z, dwb = z;
            if (ja >= 0) side_delta(ja, rdir(j), arm_a(kj), 1.0, &dva, &dwa);
            if (jb >= 0) side_delta(jb, -rdir(j), arm_b(kj), 1.0, &dvb, &dwb);
            for (int i = 0; i < nr; ++i) {
                const int ki = i >> 2, a = ba(ki), b = bb(ki);
                const bool aa = a >= 0 && a == ja, ab = a >= 0 && a == jb;
                const bool ba_ = b >= 0 && b == ja, bb_ = b >= 0 && b == jb;
                const V3<double> av = aa ? dva : ab ? dvb : z, aw = aa ? dwa : ab ? dwb : z;
                const V3<double> bv = ba_ ? dva : bb_ ? dvb : z, bw = ba_ ? dwa : bb_ ? dwb : z;
                w.ws[w.L.A + i * cap + j] = residual_of(rdir(i), a, b, arm_a(ki), arm_b(ki), av, aw, bv, bw);
            }
        }
        for (int i = 0; i < nr; ++i) {
            w.dw[w.L.rdiag + i] = w.ws[w.L.A + i * cap + i] + rreg(i);
            w.dw[w.L.rbias + i] = rbias(i);
            w.dw[w.L.f + i] = 0.0;
        }
        if (!w.newton_contacts(nr, false)) return 4;  // NEWT_ERR_NEWTON_FAIL
        for (int i = 0; i < nr; ++i) rimp(i) = w.dw[w.L.f + i];
        return 0;
    }

    // The free-body solve: per-body wrenches into bwr; 0 or a fault code.
    NEWT_HD int solve(int nc) const {
        const int nb = w.m.n_bodies;
        const double dt = w.m.dt;
        for (int i = 0; i < 6 * nb; ++i) w.dw[w.L.bwr + i] = 0.0;
        if (nc == 0 || dt <= 0.0) return 0;
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
            // a link side puts the contact in the tree solve (world.rs:2107-2115)
            if (World<double>::on_link(ga) || World<double>::on_link(gb)) continue;
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
        if (nblk == 0) return 0;
        if (w.m.mode == NEWT_SOLVER_NEWTON && w.m.n_links > 0) {
            const int st = newton_impulses(4 * nblk);
            if (st != 0) return st;
        } else {
            for (int i = 0; i < 6 * nb; ++i) w.dw[w.L.bbd + i] = 0.0;
            for (int it = 0; it < w.m.iterations; ++it)
                for (int k = 0; k < nblk; ++k) {
                    step_non_negative(4 * k);
                    step_pair(4 * k, w.dw[w.L.bblk + 8 * k + 6]);
                    step_pair(4 * k + 2, w.dw[w.L.bblk + 8 * k + 7]);
                }
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
        return 0;
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

    // ---------------- penalty contacts (world.rs:2281-2301, 2868-2949) -------
    // apply_contact_wrench of stored contact c at the state in qpos/qvel,
    // into bwr: the normal force max(0, k pen_eff - c v_n) with (k, c) =
    // solref_to_kc (geom.rs:137-145) of the combined SolRef on the reduced
    // mass; per tangent_basis axis the damping force -c v_t clamped to
    // +-mu f_n; force and r x force on each body, opposite on the second.
    NEWT_HD void penalty_wrench(int c) const {
        const int o = w.L.cgeo + 9 * c;
        const GeomT& ga = w.m.geoms[w.iw[w.L.cids + 3 * c]];
        const GeomT& gb = w.m.geoms[w.iw[w.L.cids + 3 * c + 1]];
        const int a = ga.body, b = gb.body;
        double m_eff;
        if (a >= 0 && b >= 0) {
            const double ma = body(a).mass, mb = body(b).mass;
            m_eff = ma * mb / (ma + mb);
        } else if (a >= 0) {
            m_eff = body(a).mass;
        } else if (b >= 0) {
            m_eff = body(b).mass;
        } else {
            return;
        }
        // combine_solref (geom.rs:156-185)
        double sr0 = ga.solref[0] <= gb.solref[0] ? ga.solref[0] : gb.solref[0];
        double sr1 = ga.solref[1] <= gb.solref[1] ? ga.solref[1] : gb.solref[1];
        if ((ga.solref[0] < 0.0) != (gb.solref[0] < 0.0)) {
            const GeomT& gd = ga.solref[0] < 0.0 ? ga : gb;
            sr0 = gd.solref[0];
            sr1 = gd.solref[1];
        }
        double k, cd;
        if (sr0 < 0.0) {
            k = -sr0;
            cd = -sr1;
        } else {
            k = m_eff / (sr0 * sr0);
            cd = 2.0 * sr1 * m_eff / sr0;
        }
        const V3<double> cp = w.kld3(o), n = w.kld3(o + 3);
        // point_velocity (world.rs:3038-3048)
        V3<double> ra = z3<double>(), rb = z3<double>(), va = z3<double>(), vb = z3<double>();
        if (a >= 0) {
            ra = cp - bpos(a);
            va = bvel(a) + cross3(qrot(bori(a), bomega(a)), ra);
        }
        if (b >= 0) {
            rb = cp - bpos(b);
            vb = bvel(b) + cross3(qrot(bori(b), bomega(b)), rb);
        }
        const V3<double> vrel = va - vb;
        const double vn = dot3(vrel, n);
        const double pen_eff = w.dw[o + 6] - w.dw[o + 8];
        if (pen_eff <= 0.0) return;
        const double f_raw = k * pen_eff - cd * vn;
        const double fn = f_raw > 0.0 ? f_raw : 0.0;
        if (fn <= 0.0) return;
        const V3<double> ref = nabs(n.z) < 0.9 ? mk3(0.0, 0.0, 1.0) : mk3(1.0, 0.0, 0.0);
        const V3<double> t1 = unit3(cross3(ref, n));
        const V3<double> t2 = cross3(n, t1);
        const V3<double> vt = vrel - n * vn;
        const double vt1 = dot3(vt, t1), vt2 = dot3(vt, t2);
        const double cap = w.dw[o + 7] * fn;
        const double x1 = -cd * vt1, x2 = -cd * vt2;
        const double ft1 = x1 > cap ? cap : (x1 < -cap ? -cap : x1);
        const double ft2 = x2 > cap ? cap : (x2 < -cap ? -cap : x2);
        const V3<double> fa = (n * fn + t1 * ft1) + t2 * ft2;
        if (a >= 0) {
            const int q = w.L.bwr + 6 * a;
            w.kst3(q, w.kld3(q) + fa);
            w.kst3(q + 3, (w.kld3(q + 3) + cross3(ra, fa)) + z3<double>());  // no rolling friction
        }
        if (b >= 0) {
            const V3<double> fb = -fa;
            const int q = w.L.bwr + 6 * b;
            w.kst3(q, w.kld3(q) + fb);
            w.kst3(q + 3, (w.kld3(q + 3) + cross3(rb, fb)) + z3<double>());
        }
    }

    // compute_penalty_wrenches at the state in qpos/qvel: geom poses, the
    // legacy narrowphase over the step's pairs, wrenches into bwr.  Returns
    // the contact count (above cap_c: a capacity fault, nothing applied).
    NEWT_HD int penalty_wrenches(int np) const {
        set_poses();
        const int nc = w.narrowphase<true>(np);
        for (int i = 0; i < 6 * w.m.n_bodies; ++i) w.dw[w.L.bwr + i] = 0.0;
        if (nc > w.m.cap_c) return nc;
        for (int c = 0; c < nc; ++c) penalty_wrench(c);
        return nc;
    }

    // step_bodies / step_bodies_euler under the penalty solver
    // (world.rs:1705-1785, 1790-1817): contacts and wrenches are sampled
    // again at every RK4 stage.  Stage states are written into qpos/qvel;
    // s0 and the running RK4 sum (k1 + 2 k2 + 2 k3 + k4, summed in that
    // order) sit in the row scratch, 26 doubles per body.
    NEWT_HD int step_penalty(int np) const {
        const int nb = w.m.n_bodies;
        const double dt = w.m.dt;
        int nc = penalty_wrenches(np);
        if (nc > w.m.cap_c) return 8;  // NEWT_ERR_CAPACITY
        w.iw[w.L.iwf] = np;
        const bool rk4 = w.m.integrator == NEWT_INTEGRATOR_RK4;
        const int S = w.L.brow;
        auto put = [&](int b, const St& s) {
            const BodyT& bt = body(b);
            const int qo = bt.q_off, vo = bt.v_off;
            w.qpos[qo] = s.p.x; w.qpos[qo + 1] = s.p.y; w.qpos[qo + 2] = s.p.z;
            w.qpos[qo + 3] = s.q.x; w.qpos[qo + 4] = s.q.y; w.qpos[qo + 5] = s.q.z; w.qpos[qo + 6] = s.q.w;
            w.qvel[vo] = s.v.x; w.qvel[vo + 1] = s.v.y; w.qvel[vo + 2] = s.v.z;
            w.qvel[vo + 3] = s.om.x; w.qvel[vo + 4] = s.om.y; w.qvel[vo + 5] = s.om.z;
        };
        auto save = [&](int off, V3<double> a, V3<double> b2, Qt<double> q, V3<double> c) {
            w.kst3(off, a);
            w.kst3(off + 3, b2);
            w.dw[off + 6] = q.x; w.dw[off + 7] = q.y; w.dw[off + 8] = q.z; w.dw[off + 9] = q.w;
            w.kst3(off + 10, c);
        };
        auto ldq = [&](int off) { return Qt<double>{w.dw[off], w.dw[off + 1], w.dw[off + 2], w.dw[off + 3]}; };
        for (int b = 0; b < nb; ++b) {
            const BodyT& bt = body(b);
            const St s0{bpos(b), bvel(b), bori(b), bomega(b)};
            const Dv k1 = evaluate(b, s0, w.kld3(w.L.bwr + 6 * b), w.kld3(w.L.bwr + 6 * b + 3));
            const int vo = bt.v_off;
            w.qacc[vo] = k1.dv.x; w.qacc[vo + 1] = k1.dv.y; w.qacc[vo + 2] = k1.dv.z;
            w.qacc[vo + 3] = k1.dw.x; w.qacc[vo + 4] = k1.dw.y; w.qacc[vo + 5] = k1.dw.z;
            if (rk4) {
                save(S + 26 * b, s0.p, s0.v, s0.q, s0.om);
                save(S + 26 * b + 13, k1.dp, k1.dv, k1.dq, k1.dw);
                put(b, advance(s0, k1, dt * 0.5));
            } else {
                const V3<double> v = s0.v + k1.dv * dt;
                const V3<double> om = s0.om + k1.dw * dt;
                St n;
                n.p = s0.p + v * dt;
                n.q = qintegrate(s0.q, om, dt);
                n.v = clampv(v, bt.has_max_lin, bt.max_lin);
                n.om = clampv(om, bt.has_max_ang, bt.max_ang);
                put(b, n);
            }
        }
        if (rk4) {
            for (int stg = 2; stg <= 4; ++stg) {
                nc = penalty_wrenches(np);
                if (nc > w.m.cap_c) {
                    // A world frozen here reports its start-of-step state,
                    // not the stage state written for this evaluation.
                    for (int b = 0; b < nb; ++b) {
                        const int o0 = S + 26 * b;
                        put(b, St{w.kld3(o0), w.kld3(o0 + 3), ldq(o0 + 6), w.kld3(o0 + 10)});
                    }
                    return 8;  // NEWT_ERR_CAPACITY
                }
                const double wgt = stg == 4 ? 1.0 : 2.0;
                for (int b = 0; b < nb; ++b) {
                    const BodyT& bt = body(b);
                    const int o0 = S + 26 * b, oa = o0 + 13;
                    const St s0{w.kld3(o0), w.kld3(o0 + 3), ldq(o0 + 6), w.kld3(o0 + 10)};
                    const St sc{bpos(b), bvel(b), bori(b), bomega(b)};
                    const Dv kk = evaluate(b, sc, w.kld3(w.L.bwr + 6 * b), w.kld3(w.L.bwr + 6 * b + 3));
                    const V3<double> sdp = stg == 4 ? w.kld3(oa) + kk.dp : w.kld3(oa) + kk.dp * wgt;
                    const V3<double> sdv = stg == 4 ? w.kld3(oa + 3) + kk.dv : w.kld3(oa + 3) + kk.dv * wgt;
                    const Qt<double> sdq = stg == 4 ? qadd(ldq(oa + 6), kk.dq) : qadd(ldq(oa + 6), qsc(kk.dq, wgt));
                    const V3<double> sdw = stg == 4 ? w.kld3(oa + 10) + kk.dw : w.kld3(oa + 10) + kk.dw * wgt;
                    if (stg < 4) {
                        save(oa, sdp, sdv, sdq, sdw);
                        put(b, advance(s0, kk, stg == 2 ? dt * 0.5 : dt));
                    } else {
                        const double sixth = 1.0 / 6.0;
                        St n;
                        n.p = s0.p + (sdp * sixth) * dt;
                        n.v = clampv(s0.v + (sdv * sixth) * dt, bt.has_max_lin, bt.max_lin);
                        n.q = qnormalize(qadd(s0.q, qsc(qsc(sdq, sixth), dt)));
                        n.om = clampv(s0.om + (sdw * sixth) * dt, bt.has_max_ang, bt.max_ang);
                        put(b, n);
                    }
                }
            }
        }
        w.iw[w.L.iwf + 1] = nc;
        set_poses();
        w.populate(w.m.dt, false);
        return 0;
    }

    // World::step for a model with trees and free bodies (world.rs:1459-1537),
    // PGS, or Newton where every contact has a link side: the step's pairs
    // and contacts, the tree solve with the bodies as components of the
    // tree-body rows (World::pre), the trees' integration (World::post),
    // then the free-body solve over the contacts with no link side, the
    // tree solve's body wrenches added to its own (world.rs:2097-2149), the
    // bodies' integration, and the broadphase refresh.  Trees and bodies
    // integrate independent parts of the state, so newt's order (trees
    // first under Euler, bodies first under RK4) does not change a bit.
    NEWT_HD int step_with_trees() const {
        // The tree solve's body wrenches: zero unless some contact row
        // touches a body this step.
        for (int i = 0; i < 6 * w.m.n_bodies; ++i) w.dw[w.L.tbw + i] = 0.0;
        int st = w.pre();
        if (st != 0) return st;
        st = w.post(false);
        if (st != 0) return st;
        const int nc = w.iw[w.L.iwf + w.m.n_links + 1];
        st = solve(nc);
        if (st != 0) return st;
        for (int b = 0; b < w.m.n_bodies; ++b)
            for (int k = 0; k < 6; ++k) w.dw[w.L.bwr + 6 * b + k] = w.dw[w.L.bwr + 6 * b + k] + w.dw[w.L.tbw + 6 * b + k];
        integrate();
        set_poses();
        w.forward_kinematics();
        w.link_velocities();
        w.populate(w.m.dt, false);
        return 0;
    }

    // World::step for a free-body model; 0 or a newt_status fault code.
    // Penalty selects the penalty-contact step (its own kernel instance, so
    // its registers do not weigh on the constraint-solver step).
    template <bool Penalty> NEWT_HD int step() const {
        const int np = w.active_pairs();
        if constexpr (Penalty) return step_penalty(np);
        int nc = 0;
        if (w.m.mode == NEWT_SOLVER_PGS) {
            nc = w.narrowphase(np);
            if (nc > w.m.cap_c) return 8;  // NEWT_ERR_CAPACITY
        }
        w.iw[w.L.iwf] = np;
        w.iw[w.L.iwf + 1] = nc;
        const int st = solve(nc);
        if (st != 0) return st;
        integrate();
        set_poses();
        w.populate(w.m.dt, false);
        return 0;
    }
};

}  // namespace newt

// ---------------------------------------------------------------------------
// The same step with one wavefront per world.  The body state, world inverse
// inertias, impulse accumulators, wrenches and fat bounds of the world live
// in LDS; so do the contact rows when the world's active contacts fit the
// LDS row capacity, else they go to a per-world region of global memory (the
// same code through a flat pointer).  Candidate pairs, narrowphase, row
// assembly, wrench accumulation, integration and the broadphase refresh run
// one lane per pair / contact / body / geom; the Gauss-Seidel sweeps run on
// lane 0 in the reference's order.  Every value is computed by the same
// operations as FreeBodies above, so both match the FP64 oracle bit for bit.
// ---------------------------------------------------------------------------
namespace newt {

// Offsets (elements) of one world's LDS block and its global fallback.
struct BodyWaveLayout {
    int st, iiw, bd, wr, fat, rows, blk;  // doubles
    int scr;                              // 64 doubles: one value per lane, for broadcasts
    int ints_at;                          // doubles before the int region
    int bbi, misc;                        // ints, after the doubles
    int lds_bytes;                        // bytes per world
    int lds_contacts;                     // active contacts whose rows fit in LDS
    int g_rows, g_blk, g_stride_d;        // global fallback, doubles per world
    int g_bbi, g_stride_i;                // ints per world
    // Newton solve (global, per world): rows cap nn, then H, L and Lᵀ (nn x nn), the
    // vectors lin, x, g, d, y, p, pl, pr, slope, hv, four probe points and their
    // products with H, and H times the best probe (nn each), and the break /
    // candidate lists (2 nn + 4 each)
    int nn, g_H, g_L, g_Lt, g_vec, g_brk, g_cand;  // the factor row- and column-major
};

// Rows of one world's Newton system at most (2048 = 512 contacts): H and its
// factor in both orders take 3 x 32 MiB per world at that size (3 x (4 cap_c)^2
// doubles below it).  A world with more active rows faults with NEWT_ERR_CAPACITY.
constexpr int kNewtonMaxRows = 2048;

typedef double f64x4 __attribute__((ext_vector_type(4)));
// Doubles in global memory.  The Newton workspace is addressed through these:
// from the workspace's generic pointers the compiler emits flat loads, whose
// completion waits also cover LDS traffic and serialize every load.
typedef __attribute__((address_space(1))) double gf64;

// Lane r's 64-bit value (0 for r >= 64).
__device__ __forceinline__ unsigned long long row_mask(unsigned long long v, int r) {
    if (r >= 64) return 0ull;
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, r), hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), r);
    return ((unsigned long long)hi << 32) | lo;
}

inline BodyWaveLayout make_body_wave_layout(const ModelT& m, int lds_budget_bytes) {
    BodyWaveLayout B{};
    const int nb = m.n_bodies, ng = m.n_geoms;
    const int fixed_d = 34 * nb + 6 * ng + 64;
    int lc = (lds_budget_bytes - 8 * fixed_d - 4 * 8) / (8 * 36 + 4 * 2);
    lc = std::max(0, std::min(lc, m.cap_c));
    int d = 0;
    auto D = [&](int n) { int o = d; d += n; return o; };
    B.st = D(13 * nb);
    B.iiw = D(9 * nb);
    B.bd = D(6 * nb);
    B.wr = D(6 * nb);
    B.fat = D(6 * ng);
    B.scr = D(64);
    B.rows = D(28 * lc);
    B.blk = D(8 * lc);
    B.ints_at = d;
    int i = 0;
    auto I = [&](int n) { int o = i; i += n; return o; };
    B.bbi = I(2 * lc);
    B.misc = I(8);
    B.lds_contacts = lc;
    B.lds_bytes = 8 * d + 4 * i;
    B.g_rows = 0;
    B.g_blk = 28 * m.cap_c;
    B.g_stride_d = 36 * m.cap_c;
    B.g_bbi = 0;
    B.g_stride_i = 2 * m.cap_c;
    if (m.mode == NEWT_SOLVER_NEWTON) {
        const int nn = std::min(4 * m.cap_c, kNewtonMaxRows);
        B.nn = nn;
        B.g_H = B.g_stride_d;
        B.g_L = B.g_H + nn * nn;
        B.g_Lt = B.g_L + nn * nn;
        B.g_vec = B.g_Lt + nn * nn;
        B.g_brk = B.g_vec + 19 * nn;
        B.g_cand = B.g_brk + 2 * nn + 4;
        B.g_stride_d = B.g_cand + 2 * nn + 4;
    }
    return B;
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// The Gauss-Seidel sweeps of solve_free_bodies (solver.rs:1066-1170) over
// nblk pyramidal condim-3 blocks, on one lane: pgs_step_non_negative and two
// pgs_step_pyramidal_pair per block, body impulse accumulators in bd.  Row
// ri: dir, reg, diag, bias, impulse (7 doubles); block k: arm_a, arm_b, h01,
// h23 (8 doubles) and its bodies in bbi.  Every pointer is a parameter so
// the compiler keeps them in registers (the regions do not overlap).
__device__ __noinline__ void body_pgs_sweeps(int iterations, int nblk, double* __restrict__ rows,
                                             const double* __restrict__ blk, const int* __restrict__ bbi,
                                             double* __restrict__ bd, const double* __restrict__ iiw,
                                             const double* __restrict__ st, const BodyT* __restrict__ bodies) {
    // One block's rows, arms and bodies in registers: every load of a block
    // is issued together (the rows may sit in global memory), and the four
    // impulses stay in registers until the block is done.
    struct Blk {
        V3<double> dir[4];
        double reg[4], diag[4], bias[4], f[4];
        V3<double> arma, armb;
        double h01, h23;
        int a, b;
    };
    auto load = [&](int k, Blk& q) {
        const double* r = rows + 28 * k;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            q.dir[u] = mk3(r[7 * u], r[7 * u + 1], r[7 * u + 2]);
            q.reg[u] = r[7 * u + 3];
            q.diag[u] = r[7 * u + 4];
            q.bias[u] = r[7 * u + 5];
            q.f[u] = r[7 * u + 6];
        }
        const double* p = blk + 8 * k;
        q.arma = mk3(p[0], p[1], p[2]);
        q.armb = mk3(p[3], p[4], p[5]);
        q.h01 = p[6];
        q.h23 = p[7];
        q.a = bbi[2 * k];
        q.b = bbi[2 * k + 1];
    };
    auto ld3 = [](const double* p) { return mk3(p[0], p[1], p[2]); };
    auto ori = [&](int b) {
        const double* p = st + 13 * b + 3;
        return Qt<double>{p[0], p[1], p[2], p[3]};
    };
    auto residual = [&](const Blk& q, int u) {
        double s = 0.0;
        if (q.a >= 0) {
            V3<double> wa = qrot(ori(q.a), ld3(bd + 6 * q.a + 3));
            V3<double> vp = ld3(bd + 6 * q.a) + cross3(wa, q.arma);
            s += dot3(vp, q.dir[u]);
        }
        if (q.b >= 0) {
            V3<double> wb = qrot(ori(q.b), ld3(bd + 6 * q.b + 3));
            V3<double> vp = ld3(bd + 6 * q.b) + cross3(wb, q.armb);
            s -= dot3(vp, q.dir[u]);
        }
        return s;
    };
    auto side = [&](int i, V3<double> dir, V3<double> arm, double df) {
        double* p = bd + 6 * i;
        V3<double> f = dir * df;
        V3<double> tq = cross3(arm, f);
        M33<double> w3;
        for (int k = 0; k < 9; ++k) w3.m[k] = iiw[9 * i + k];
        V3<double> dv = ld3(p) + f / bodies[i].mass;
        V3<double> dw = ld3(p + 3) + qrot_inv(ori(i), mv(w3, tq));
        p[0] = dv.x; p[1] = dv.y; p[2] = dv.z;
        p[3] = dw.x; p[4] = dw.y; p[5] = dw.z;
    };
    auto apply = [&](const Blk& q, int u, double df) {
        if (df == 0.0) return;
        if (q.a >= 0) side(q.a, q.dir[u], q.arma, df);
        if (q.b >= 0) side(q.b, -q.dir[u], q.armb, df);
    };
    auto non_negative = [&](Blk& q, int u) {
        const double f = q.f[u];
        const double res = residual(q, u) + q.reg[u] * f + q.bias[u];
        double delta = -res / q.diag[u];
        const double nf = f + delta;
        const double pr = nf < 0.0 ? 0.0 : nf;
        delta = pr - f;
        q.f[u] = pr;
        apply(q, u, delta);
    };
    auto pair = [&](Blk& q, int u, double h01) {
        const int v = u + 1;
        const double fi = q.f[u], fj = q.f[v];
        const double res_i = residual(q, u) + q.reg[u] * fi + q.bias[u];
        const double res_j = residual(q, v) + q.reg[v] * fj + q.bias[v];
        const double h00 = q.diag[u], h11 = q.diag[v];
        const double det = h00 * h11 - h01 * h01;
        if (det <= 0.0) {
            non_negative(q, u);
            non_negative(q, v);
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
        q.f[u] = a;
        q.f[v] = b;
        apply(q, u, a - fi);
        apply(q, v, b - fj);
    };
    for (int it = 0; it < iterations; ++it) {
        Blk q;
        if (nblk > 0) load(0, q);
        for (int k = 0; k < nblk; ++k) {
            Blk nx;
            if (k + 1 < nblk) load(k + 1, nx);  // issued before this block's chain
            non_negative(q, 0);
            pair(q, 0, q.h01);
            pair(q, 2, q.h23);
            double* r = rows + 28 * k;
#pragma unroll
            for (int u = 0; u < 4; ++u) r[7 * u + 6] = q.f[u];
            q = nx;
        }
    }
}

struct BodyWave {
    const ModelT& m;
    const BodyWaveLayout& B;
    const Layout& L;
    const DataT& d;
    double* dws;      // lane-interleaved contact geometry (Layout::cgeo)
    int* iws;         // lane-interleaved contact ids and pair list
    double* sd;       // this world's LDS doubles
    int* si;          // this world's LDS ints
    double* gd;       // this world's global fallback (doubles)
    int* gi;          // and ints
    int world, lane;

    // rows / blocks: LDS or global, chosen per step
    mutable double* rows;
    mutable double* blk;
    mutable int* bbi;
    // 1: newt's operation order everywhere (FP64_EXACT); 0: the Newton back
    // substitution runs column by column (the drop-in's 1e-9 contract).
    int exact;

    __device__ __forceinline__ double& cg(int k, int j) const { return dws[(size_t)(L.cgeo + 9 * k + j) * d.W + world]; }
    __device__ __forceinline__ int& cid(int k, int j) const { return iws[(size_t)(L.cids + 3 * k + j) * d.W + world]; }

    __device__ __forceinline__ const BodyT& body(int b) const { return m.bodies[b]; }
    __device__ __forceinline__ V3<double> ld3(const double* p) const { return mk3(p[0], p[1], p[2]); }
    __device__ __forceinline__ void st3(double* p, V3<double> v) const { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
    __device__ __forceinline__ V3<double> bpos(int b) const { return ld3(sd + B.st + 13 * b); }
    __device__ __forceinline__ Qt<double> bori(int b) const {
        const double* p = sd + B.st + 13 * b + 3;
        return Qt<double>{p[0], p[1], p[2], p[3]};
    }
    __device__ __forceinline__ V3<double> bvel(int b) const { return ld3(sd + B.st + 13 * b + 7); }
    __device__ __forceinline__ V3<double> bomega(int b) const { return ld3(sd + B.st + 13 * b + 10); }
    __device__ __forceinline__ M33<double> iiw(int b) const {
        M33<double> a;
        for (int k = 0; k < 9; ++k) a.m[k] = sd[B.iiw + 9 * b + k];
        return a;
    }
    __device__ __forceinline__ V3<double> gravity_of(int b) const {
        const V3<double> g = mk3(m.gravity[0], m.gravity[1], m.gravity[2]);
        const double s = body(b).gravity_scale;
        return s == 1.0 ? g : g * s;
    }

    // ---------------- state in and out ------------------------------------
    __device__ __forceinline__ void load() const {
        const size_t W = d.W;
        for (int b = lane; b < m.n_bodies; b += 64) {
            const BodyT& bt = body(b);
            double* s = sd + B.st + 13 * b;
            for (int k = 0; k < 7; ++k) s[k] = d.qpos[(size_t)(bt.q_off + k) * W + world];
            for (int k = 0; k < 6; ++k) s[7 + k] = d.qvel[(size_t)(bt.v_off + k) * W + world];
        }
        for (int k = lane; k < 6 * m.n_geoms; k += 64) sd[B.fat + k] = d.fat[(size_t)k * W + world];
        gsync();
    }
    __device__ __forceinline__ void store() const {
        const size_t W = d.W;
        for (int b = lane; b < m.n_bodies; b += 64) {
            const BodyT& bt = body(b);
            const double* s = sd + B.st + 13 * b;
            for (int k = 0; k < 7; ++k) d.qpos[(size_t)(bt.q_off + k) * W + world] = s[k];
            for (int k = 0; k < 6; ++k) d.qvel[(size_t)(bt.v_off + k) * W + world] = s[7 + k];
        }
        for (int k = lane; k < 6 * m.n_geoms; k += 64) d.fat[(size_t)k * W + world] = sd[B.fat + k];
    }

    // ---------------- broadphase (world.rs:1254-1358) ----------------------
    __device__ __forceinline__ void geom_pose(const GeomT& g, V3<double>* p, Qt<double>* o) const {
        V3<double> gp = mk3(g.pos[0], g.pos[1], g.pos[2]);
        Qt<double> gq{g.quat[0], g.quat[1], g.quat[2], g.quat[3]};
        if (g.body < 0) {
            *p = gp;
            *o = gq;
            return;
        }
        Qt<double> bo = bori(g.body);
        *p = bpos(g.body) + qrot(bo, gp);
        *o = qmul(bo, gq);
    }
    __device__ __forceinline__ void populate(bool insert_all) const {
        const double dt = m.dt;
        for (int gi0 = lane; gi0 < m.n_geoms; gi0 += 64) {
            const GeomT& g = m.geoms[gi0];
            V3<double> sp, smn, smx, emn, emx;
            Qt<double> so;
            geom_pose(g, &sp, &so);
            World<double>::geom_aabb_of(g, sp, so, &smn, &smx);
            double ang_speed = 0.0;
            V3<double> pivot = sp;
            if (g.body >= 0) {
                const V3<double> bp = bpos(g.body), bv = bvel(g.body), bw = bomega(g.body);
                const Qt<double> bo = bori(g.body);
                V3<double> ep = bp + bv * dt;
                Qt<double> eo = qintegrate(bo, bw, dt);
                V3<double> gp = mk3(g.pos[0], g.pos[1], g.pos[2]);
                Qt<double> gq{g.quat[0], g.quat[1], g.quat[2], g.quat[3]};
                World<double>::geom_aabb_of(g, ep + qrot(eo, gp), qmul(eo, gq), &emn, &emx);
                ang_speed = len3(qrot(bo, bw));
                pivot = bp;
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
            V3<double> s3 = mk3(sweep, sweep, sweep);
            mn = mn - s3;
            mx = mx + s3;
            V3<double> m3 = mk3(g.margin, g.margin, g.margin);
            mn = mn - m3;
            mx = mx + m3;
            double* f = sd + B.fat + 6 * gi0;
            bool inside = f[0] <= mn.x && f[1] <= mn.y && f[2] <= mn.z && f[3] >= mx.x && f[4] >= mx.y && f[5] >= mx.z;
            if (!inside || insert_all) {
                V3<double> e = mx - mn;
                V3<double> fm = mk3(nmax(e.x * 0.1, 1.0e-4), nmax(e.y * 0.1, 1.0e-4), nmax(e.z * 0.1, 1.0e-4));
                V3<double> a = mn - fm, b = mx + fm;
                f[0] = a.x; f[1] = a.y; f[2] = a.z;
                f[3] = b.x; f[4] = b.y; f[5] = b.z;
            }
        }
        gsync();
    }

    // ---------------- pairs + narrowphase (contact.rs:269-403, 569-1196) --
    // Contacts of candidate pair k into (cp, nrm, pen); -1 when the fat
    // bounds do not overlap.  Same operations as World::narrowphase.
    __device__ __forceinline__ int pair_contacts(int k, V3<double>* cp, V3<double>* nrm, double* pen, double* fr_out,
                                                 double* gap_out) const {
        const PairT& pr = m.pairs[k];
        const double* fa = sd + B.fat + 6 * pr.a;
        const double* fb = sd + B.fat + 6 * pr.b;
        bool ov = fa[0] <= fb[3] && fa[3] >= fb[0] && fa[1] <= fb[4] && fa[4] >= fb[1] && fa[2] <= fb[5] &&
                  fa[5] >= fb[2];
        if (!ov) return -1;
        const GeomT& ga = m.geoms[pr.a];
        const GeomT& gb = m.geoms[pr.b];
        *fr_out = ga.friction < gb.friction ? ga.friction : gb.friction;
        const double margin = ga.margin > gb.margin ? ga.margin : gb.margin;
        *gap_out = ga.gap > gb.gap ? ga.gap : gb.gap;
        if (m.slop > 0.0 && *gap_out < m.slop) *gap_out = m.slop;  // append_contacts (world.rs:2517-2528)
        if (ga.shape == SH_BOX && gb.shape == SH_BOX) {
            V3<double> pa, pb, n;
            Qt<double> qa, qb;
            geom_pose(ga, &pa, &qa);
            geom_pose(gb, &pb, &qb);
            const int c = box_box(pa, qa, ga.size, pb, qb, gb.size, margin, cp, pen, &n);
            for (int i = 0; i < c; ++i) nrm[i] = n;
            return c;
        }
        if (ga.shape != SH_PLANE && gb.shape != SH_PLANE) {
            V3<double> pa, pb;
            Qt<double> qa, qb;
            geom_pose(ga, &pa, &qa);
            geom_pose(gb, &pb, &qb);
            return round_pair(ga.shape == SH_CAPSULE, pa, qa, ga.size, gb.shape == SH_CAPSULE, pb, qb, gb.size, margin,
                              &cp[0], &nrm[0], &pen[0])
                       ? 1
                       : 0;
        }
        const bool flip = gb.shape != SH_PLANE;
        const GeomT& gp = flip ? gb : ga;
        const GeomT& gq = flip ? ga : gb;
        V3<double> pp, qp;
        Qt<double> po, qo;
        geom_pose(gq, &pp, &po);
        geom_pose(gp, &qp, &qo);
        const V3<double> n = qrot(po, mk3(0.0, 0.0, 1.0));
        const V3<double> nn = flip ? -n : n;
        int cnt = 0;
        if (gp.shape == SH_SPHERE) {
            const double radius = gp.size[0];
            const double raw = dot3(qp - pp, n) - radius;
            if (raw <= margin) {
                cp[cnt] = qp - n * (radius + raw * 0.5);
                nrm[cnt] = nn;
                pen[cnt++] = margin - raw;
            }
        } else if (gp.shape == SH_BOX) {
            const double dist = dot3(qp - pp, n);
            for (int i = 0; i < 8; ++i) {
                V3<double> loc = mk3((i & 1) == 0 ? -gp.size[0] : gp.size[0], (i & 2) == 0 ? -gp.size[1] : gp.size[1],
                                     (i & 4) == 0 ? -gp.size[2] : gp.size[2]);
                V3<double> co = qrot(qo, loc);
                double ld = dot3(n, co);
                double raw = dist + ld;
                if (raw > margin || ld > 0.0) continue;
                if (cnt >= 4) continue;
                cp[cnt] = (qp + co) - n * (raw * 0.5);
                nrm[cnt] = nn;
                pen[cnt++] = margin - raw;
            }
        } else if (gp.shape == SH_CAPSULE) {
            const double radius = gp.size[0], hh = gp.size[1];
            const V3<double> aw = qrot(qo, mk3(0.0, 0.0, 1.0));
            for (int e = 0; e < 2; ++e) {
                V3<double> end = e == 0 ? qp + aw * hh : qp - aw * hh;
                double raw = dot3(end - pp, n) - radius;
                if (raw <= margin) {
                    cp[cnt] = end - n * (radius + raw * 0.5);
                    nrm[cnt] = nn;
                    pen[cnt++] = margin - raw;
                }
            }
        }
        return cnt;
    }
    // Contacts of this step in pair order into Layout::cids/cgeo; active
    // pair indices into Layout::pairs.  Returns the contact count (cap_c + 1
    // when the world holds more than the compiled capacity).
    __device__ __forceinline__ int contacts(int* np_out) const {
        int nc = 0, np = 0;
        bool over = false;
        for (int base = 0; base < m.n_pairs; base += 64) {
            const int k = base + lane;
            V3<double> cp[4], nrm[4];
            double pen[4], fr = 0.0, gap = 0.0;
            int c = -1;
            if (k < m.n_pairs) c = pair_contacts(k, cp, nrm, pen, &fr, &gap);
            const unsigned long long act = __ballot(c >= 0);
            if (c >= 0) iws[(size_t)(L.pairs + np + __popcll(act & lanes_below(lane))) * d.W + world] = k;
            np += __popcll(act);
            const int cc = c > 0 ? c : 0;
            const int incl = wave_incl_scan(cc, lane);
            const int total = __shfl(incl, 63, 64);
            const int off = nc + incl - cc;
            if (off + cc <= m.cap_c) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (i >= cc) break;
                    const int o = off + i;
                    cid(o, 0) = m.pairs[k].a;
                    cid(o, 1) = m.pairs[k].b;
                    cid(o, 2) = i;
                    cg(o, 0) = cp[i].x; cg(o, 1) = cp[i].y; cg(o, 2) = cp[i].z;
                    cg(o, 3) = nrm[i].x; cg(o, 4) = nrm[i].y; cg(o, 5) = nrm[i].z;
                    cg(o, 6) = pen[i];
                    cg(o, 7) = fr;
                    cg(o, 8) = gap;
                }
            }
            nc += total;
            if (nc > m.cap_c) over = true;
        }
        gsync();
        *np_out = np;
        return over ? m.cap_c + 1 : nc;
    }

    // ---------------- rows (solver.rs:1892-2069, 2131-2154, 4012-4036) -----
    __device__ __forceinline__ double residual_of(V3<double> dir, int a, int b, V3<double> arma, V3<double> armb,
                                                  V3<double> dva, V3<double> dwa, V3<double> dvb, V3<double> dwb) const {
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
    // residual_of with the two bodies' orientations given.
    __device__ __forceinline__ double residual_q(V3<double> dir, int a, int b, Qt<double> qa, Qt<double> qb,
                                                 V3<double> arma, V3<double> armb, V3<double> dva, V3<double> dwa,
                                                 V3<double> dvb, V3<double> dwb) const {
        double s = 0.0;
        if (a >= 0) {
            V3<double> wa = qrot(qa, dwa);
            V3<double> vp = dva + cross3(wa, arma);
            s += dot3(vp, dir);
        }
        if (b >= 0) {
            V3<double> wb = qrot(qb, dwb);
            V3<double> vp = dvb + cross3(wb, armb);
            s -= dot3(vp, dir);
        }
        return s;
    }
    __device__ __forceinline__ void side_delta(int i, V3<double> dir, V3<double> arm, double df, V3<double>* dv,
                                               V3<double>* dwv) const {
        V3<double> f = dir * df;
        V3<double> tq = cross3(arm, f);
        *dv = *dv + f / body(i).mass;
        *dwv = *dwv + qrot_inv(bori(i), mv(iiw(i), tq));
    }
    __device__ __forceinline__ V3<double> rdir(int ri) const { return ld3(rows + 7 * ri); }

    // Block k's four rows, arms, bodies and cross responses from contact c.
    __device__ __forceinline__ void assemble(int k, int c) const {
        const double dt = m.dt;
        const GeomT& ga = m.geoms[cid(c, 0)];
        const GeomT& gb = m.geoms[cid(c, 1)];
        const int a = ga.body, b = gb.body;
        const double pen = cg(c, 6) - cg(c, 8);
        const double mu = ga.friction < gb.friction ? ga.friction : gb.friction;
        const V3<double> cp = mk3(cg(c, 0), cg(c, 1), cg(c, 2)), n = mk3(cg(c, 3), cg(c, 4), cg(c, 5));
        const V3<double> ref = nabs(n.z) < 0.9 ? mk3(0.0, 0.0, 1.0) : mk3(1.0, 0.0, 0.0);
        const V3<double> t1 = unit3(cross3(ref, n));
        const V3<double> t2 = cross3(n, t1);
        const V3<double> arma = a >= 0 ? cp - bpos(a) : z3<double>();
        const V3<double> armb = b >= 0 ? cp - bpos(b) : z3<double>();
        bbi[2 * k] = a;
        bbi[2 * k + 1] = b;
        st3(blk + 8 * k, arma);
        st3(blk + 8 * k + 3, armb);
        const V3<double> dirs[4] = {n + t1 * mu, n - t1 * mu, n + t2 * mu, n - t2 * mu};
        double sr0 = ga.solref[0] <= gb.solref[0] ? ga.solref[0] : gb.solref[0];
        double sr1 = ga.solref[1] <= gb.solref[1] ? ga.solref[1] : gb.solref[1];
        if (ga.solref[0] < 0.0 || gb.solref[0] < 0.0) {
            if (!(ga.solref[0] < 0.0 && gb.solref[0] < 0.0)) {
                const GeomT& gd = ga.solref[0] < 0.0 ? ga : gb;
                sr0 = gd.solref[0];
                sr1 = gd.solref[1];
            }
        }
        const double si4[4] = {ga.solimp[0] < gb.solimp[0] ? ga.solimp[0] : gb.solimp[0],
                               ga.solimp[1] < gb.solimp[1] ? ga.solimp[1] : gb.solimp[1],
                               ga.solimp[2] < gb.solimp[2] ? ga.solimp[2] : gb.solimp[2], ga.solimp[3]};
        const SolImpE se = solimp_effective(si4, ga.power);
        const double imp = impedance_at(se, -pen);
        for (int r = 0; r < 4; ++r) {
            double* row = rows + 7 * (4 * k + r);
            const V3<double> dir = dirs[r];
            st3(row, dir);
            double aii = 0.0;
            if (a >= 0) {
                V3<double> rc = cross3(arma, dir);
                aii += dot3(rc, mv(iiw(a), rc)) + 1.0 / body(a).mass;
            }
            if (b >= 0) {
                V3<double> rc = cross3(armb, dir);
                aii += dot3(rc, mv(iiw(b), rc)) + 1.0 / body(b).mass;
            }
            double reg = 0.0;
            if (imp > 0.0) {
                double tran = 0.0;
                if (a >= 0) tran += 1.0 / body(a).mass;
                if (b >= 0) tran += 1.0 / body(b).mass;
                const double base = (1.0 - imp) / imp * (tran * (1.0 + mu * mu));
                reg = 2.0 * mu * mu * base;
            }
            row[3] = reg;
            row[4] = aii + reg;
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
            row[5] = dvf - aref * dt;
            row[6] = 0.0;
        }
        for (int p = 0; p < 2; ++p) {
            const V3<double> applied = dirs[2 * p + 1];
            V3<double> dva = z3<double>(), dwa = z3<double>(), dvb = z3<double>(), dwb = z3<double>();
            if (a >= 0) side_delta(a, applied, arma, 1.0, &dva, &dwa);
            if (b >= 0) side_delta(b, -applied, armb, 1.0, &dvb, &dwb);
            blk[8 * k + 6 + p] = residual_of(dirs[2 * p], a, b, arma, armb, dva, dwa, dvb, dwb);
        }
    }

    // ---------------- Newton (solver.rs:1295-1351, newton.rs:275-556) ------
    // Sequential sum a_0 b_0 + a_1 b_1 + ... from -0.0 (Iterator::sum), on
    // every lane (same result everywhere).  Each lane forms one product of a
    // chunk of 64 (all loads in flight at once); the additions then run in
    // index order on the broadcast products.
    template <class PA, class PB>
    __device__ __forceinline__ double dot_seq(PA a, PB b, int n) const {
        double s = -0.0;
        for (int i0 = 0; i0 < n; i0 += 64) {
            const int i = i0 + lane;
            const double p = i < n ? a[i] * b[i] : 0.0;
            const int cnt = n - i0 < 64 ? n - i0 : 64;
            if (cnt == 64) {
#pragma unroll
                for (int u = 0; u < 64; ++u) s = s + readlane_f64(p, u);
            } else {
                for (int u = 0; u < cnt; ++u) s = s + readlane_f64(p, u);
            }
        }
        return s;
    }
    // out = H v, one lane per row, each row summed in column order from 0.0;
    // a lane carries four rows at once (independent sums, same order each),
    // and eight columns' loads are issued before their products are summed.
    template <class PH, class PV, class PO>
    __device__ __forceinline__ void mat_vec(PH H, PV v, PO out, int n) const {
        for (int i0 = lane; i0 < n; i0 += 256) {
            const int i1 = i0 + 64 < n ? i0 + 64 : i0, i2 = i0 + 128 < n ? i0 + 128 : i0, i3 = i0 + 192 < n ? i0 + 192 : i0;
            const auto h0 = H + (size_t)i0 * n, h1 = H + (size_t)i1 * n, h2 = H + (size_t)i2 * n, h3 = H + (size_t)i3 * n;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int j = 0;
            for (; j + 8 <= n; j += 8) {
                double vv[8], a0[8], a1[8], a2[8], a3[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    vv[u] = v[j + u];
                    a0[u] = h0[j + u];
                    a1[u] = h1[j + u];
                    a2[u] = h2[j + u];
                    a3[u] = h3[j + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    s0 += a0[u] * vv[u];
                    s1 += a1[u] * vv[u];
                    s2 += a2[u] * vv[u];
                    s3 += a3[u] * vv[u];
                }
            }
            for (; j < n; ++j) {
                const double vj = v[j];
                s0 += h0[j] * vj;
                s1 += h1[j] * vj;
                s2 += h2[j] * vj;
                s3 += h3[j] * vj;
            }
            out[i0] = s0;
            if (i0 + 64 < n) out[i0 + 64] = s1;
            if (i0 + 128 < n) out[i0 + 128] = s2;
            if (i0 + 192 < n) out[i0 + 192] = s3;
        }
        gsync();
    }
    // out_k = H v_k for the K <= 4 vectors v_k = vb + k nn (out_k = ob + k nn):
    // each row summed in column order from 0.0 exactly as mat_vec, so every
    // out_k equals mat_vec(H, v_k) bit for bit; one pass over H serves all K.
    template <class PH, class PV>
    __device__ __forceinline__ void mat_vec4(PH H, PV vb, PV ob, int K, int n, int nn) const {
        const int k1 = K > 1 ? 1 : 0, k2 = K > 2 ? 2 : k1, k3 = K > 3 ? 3 : k2;
        const auto v0 = vb, v1 = vb + (size_t)k1 * nn, v2 = vb + (size_t)k2 * nn, v3 = vb + (size_t)k3 * nn;
        for (int base = 0; base < n; base += 256) {
            const int i0 = base + lane;
            int ir[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) ir[r] = i0 + 64 * r < n ? i0 + 64 * r : n - 1;
            double s[4][4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int k = 0; k < 4; ++k) s[r][k] = 0.0;
            int j = 0;
            for (; j + 8 <= n; j += 8) {
                double h[4][8], vv[4][8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    vv[0][u] = v0[j + u];
                    vv[1][u] = v1[j + u];
                    vv[2][u] = v2[j + u];
                    vv[3][u] = v3[j + u];
#pragma unroll
                    for (int r = 0; r < 4; ++r) h[r][u] = H[(size_t)ir[r] * n + j + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int k = 0; k < 4; ++k) s[r][k] += h[r][u] * vv[k][u];
            }
            for (; j < n; ++j) {
                const double w0 = v0[j], w1 = v1[j], w2 = v2[j], w3 = v3[j];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double hr = H[(size_t)ir[r] * n + j];
                    s[r][0] += hr * w0;
                    s[r][1] += hr * w1;
                    s[r][2] += hr * w2;
                    s[r][3] += hr * w3;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 64 * r;
                if (i < n) {
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (k < K) ob[(size_t)k * nn + i] = s[r][k];
                }
            }
        }
        gsync();
    }
    template <class PX, class PD, class PO>
    __device__ __forceinline__ void affine(PX x, PD d, double alpha, PO out, int n) const {
        for (int i = lane; i < n; i += 64) {
            const double v = x[i] + alpha * d[i];
            out[i] = v < 0.0 ? 0.0 : v;
        }
        gsync();
    }
    // quadratic_cost 0.5 xᵀHx + linᵀx with Hx into hv.
    template <class P>
    __device__ __forceinline__ double cost(P H, P lin, P x, P hv, int n) const {
        mat_vec(H, x, hv, n);
        return 0.5 * dot_seq(x, hv, n) + dot_seq(lin, x, n);
    }
    // Impulses of rows 0..4 nblk into rows[7 ri + 6]; 0 or a fault code.
    // The substitutions keep a lane's share of a vector (rows lane + 64 c) in
    // C registers, so n <= 64 C: C = 16 up to 1024 rows, C = 32 (whose
    // arrays the compiler spills) up to 2048.
    template <int C>
    __device__ __noinline__ int newton(int nblk) const {
        static_assert(C == 16 || C == 32, "16 or 32 row chunks");
        const int n = 4 * nblk;
        if (n > B.nn || n > 64 * C) return 8;  // NEWT_ERR_CAPACITY
        const int nn = B.nn;
        gf64* H = (gf64*)(gd + B.g_H);
        gf64* Lf = (gf64*)(gd + B.g_L);   // the factor, row-major
        gf64* Lt = (gf64*)(gd + B.g_Lt);  // Lt[j n + i] = L[i][j], column-major
        gf64* V = (gf64*)(gd + B.g_vec);
        gf64 *lin = V, *x = V + nn, *g = V + 2 * nn, *dd = V + 3 * nn, *y = V + 4 * nn, *pt = V + 5 * nn;
        gf64 *pl = V + 6 * nn, *pr = V + 7 * nn, *sl = V + 8 * nn, *hv = V + 9 * nn;
        gf64* brk = (gf64*)(gd + B.g_brk);
        gf64* cand = (gf64*)(gd + B.g_cand);
        // A[i][j] = response of row i to a unit impulse on row j; H = A + diag(reg).
        // One lane per column j; rows go by contact block (four rows share
        // their bodies).  A block with neither of column j's bodies keeps
        // zero accumulators, whose residual is +0.0.  Blocks go in chunks of
        // 64: lane u loads block kc + u's row directions, arms, bodies and
        // their orientations, and the column lanes take each block from
        // broadcasts, so no load waits behind the stores of H.  The loops
        // keep every lane active (lanes past the last column work on column
        // n - 1 and store nothing), so each broadcast reads a register every
        // lane has written.
        for (int kc = 0; kc < nblk; kc += 64) {
            const int kl = kc + lane < nblk ? kc + lane : nblk - 1;
            const int la = bbi[2 * kl], lb = bbi[2 * kl + 1];
            V3<double> ldir[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) ldir[r] = rdir(4 * kl + r);
            const V3<double> lam = ld3(blk + 8 * kl), lbm = ld3(blk + 8 * kl + 3);
            const Qt<double> qid{0.0, 0.0, 0.0, 1.0};
            const Qt<double> lqa = la >= 0 ? bori(la) : qid, lqb = lb >= 0 ? bori(lb) : qid;
            const int kcnt = nblk - kc < 64 ? nblk - kc : 64;
            for (int j0 = 0; j0 < n; j0 += 64) {
                const bool on = j0 + lane < n;
                const int j = on ? j0 + lane : n - 1;
                const int kj = j >> 2;
                const int aj = bbi[2 * kj], bj = bbi[2 * kj + 1];
                const V3<double> dir = rdir(j);
                V3<double> dva = z3<double>(), dwa = z3<double>(), dvb = z3<double>(), dwb = z3<double>();
                if (aj >= 0) side_delta(aj, dir, ld3(blk + 8 * kj), 1.0, &dva, &dwa);
                if (bj >= 0) side_delta(bj, -dir, ld3(blk + 8 * kj + 3), 1.0, &dvb, &dwb);
                for (int u = 0; u < kcnt; ++u) {
                    const int ki = kc + u;
                    const int ai = __builtin_amdgcn_readlane(la, u), bi = __builtin_amdgcn_readlane(lb, u);
                    V3<double> av = z3<double>(), aw = z3<double>(), bv = z3<double>(), bw = z3<double>();
                    bool any = false;
                    if (ai >= 0 && ai == aj) { av = dva; aw = dwa; any = true; }
                    if (ai >= 0 && ai == bj) { av = dvb; aw = dwb; any = true; }
                    if (bi >= 0 && bi == aj) { bv = dva; bw = dwa; any = true; }
                    if (bi >= 0 && bi == bj) { bv = dvb; bw = dwb; any = true; }
                    double a4[4] = {0.0, 0.0, 0.0, 0.0};
                    if (__ballot(any) != 0) {
                        auto bc3 = [&](V3<double> v) {
                            return mk3(readlane_f64(v.x, u), readlane_f64(v.y, u), readlane_f64(v.z, u));
                        };
                        auto bcq = [&](Qt<double> q) {
                            return Qt<double>{readlane_f64(q.x, u), readlane_f64(q.y, u), readlane_f64(q.z, u),
                                              readlane_f64(q.w, u)};
                        };
                        const V3<double> arma = bc3(lam), armb = bc3(lbm);
                        const Qt<double> qa = bcq(lqa), qb = bcq(lqb);
                        V3<double> dr[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) dr[r] = bc3(ldir[r]);
                        if (any) {
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                a4[r] = residual_q(dr[r], ai, bi, qa, qb, arma, armb, av, aw, bv, bw);
                        }
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = 4 * ki + r;
                        double a = a4[r];
                        if (i == j) a += rows[7 * i + 3];
                        if (on) H[(size_t)i * n + j] = a;
                    }
                }
            }
        }
        for (int i = lane; i < n; i += 64) {
            lin[i] = rows[7 * i + 5];
            x[i] = 0.0;
        }
        gsync();
        // Cholesky-Banachiewicz (dynamics.rs:602-624), by panels of sixteen
        // columns.  Every entry L[i][j] receives H[i][j] minus L[i][k] L[j][k]
        // for k = 0, 1, ..., j-1 in that order, as in the row-wise loop: first
        // the terms k < j0 (left of the panel, sixteen columns accumulated at
        // once per row, so each L[i][k] load serves sixteen entries, four k
        // at a time with every load in flight together), then the in-panel
        // terms k = j0 .. j-1, column by column.  The diagonal sums
        // H[i][i] - L[i][0]^2 - L[i][1]^2 - ... are carried in `pt`, each term
        // subtracted right after its L[i][k] is computed (k increasing).  L is
        // kept column-major (Lt, read by the factorization and the forward
        // substitution) and row-major (Lf, read by the back substitution).
        constexpr int P = 16;
        const bool mfma_left = !exact;
        // Lane R holds row block R's mask of column blocks of L with a
        // nonzero entry, set as each panel is written.  A block is 16 rows
        // or columns up to 1024 rows and 32 up to 2048 (SH = log2 of the
        // block), so R < 64 and a mask fits 64 bits.  The matrix-core left
        // part visits only the column blocks where both the panel's rows and
        // the chunk's rows have one: every skipped product is an exact zero.
        constexpr int SH = C == 16 ? 4 : 5;
        unsigned long long rmask = 0;
        for (int i = lane; i < n; i += 64) pt[i] = H[(size_t)i * n + i];
        gsync();
        for (int j0 = 0; j0 < n; j0 += P) {
            const int pw = n - j0 < P ? n - j0 : P;  // panel width
            // The panel rows' own entries L[j0 + t][j0 + c] (lane t of the
            // first chunk), for the later chunks.
            double q0[P];
            // Rows j0 .. n-1, one per lane, in chunks of 64.
            for (int r0 = j0; r0 < n; r0 += 64) {
                const int i = r0 + lane;
                const bool on = i < n;
                const int ii = on ? i : n - 1;
                double acc[P];
#pragma unroll
                for (int c = 0; c < P; ++c) acc[c] = (c < pw && on) ? H[(size_t)ii * n + j0 + c] : 0.0;
                double pr = on ? pt[ii] : 0.0;  // the row's diagonal sum, with its H entries
                int k = 0;
                if (mfma_left && j0 > 0) {
                    // The left part on the matrix cores (FP64/mixed only): the
                    // 16 x 64 block sum_k L[j0 + c][k] L[i][k] as four 16 x 16
                    // tiles T of v_mfma_f64_16x16x4 (rows i = r0 + 16 T + 0..15),
                    // four k per instruction.  Lane l feeds A[l % 16][l / 16] =
                    // L[j0 + l % 16][k0 + l / 16] and B[l / 16][l % 16] =
                    // L[r0 + 16 T + l % 16][k0 + l / 16]; register r of lane l
                    // returns column c = 4 r + l / 16 of row r0 + 16 T + l % 16.
                    // Sixteen k go per group, the next group's 20 loads in
                    // flight while the current group's 16 products run.  j0 is
                    // a multiple of 16.
                    f64x4 t[4];
#pragma unroll
                    for (int T = 0; T < 4; ++T) t[T] = f64x4{0.0, 0.0, 0.0, 0.0};
                    // The panel's rows lie in block j0 >> SH, the chunk's in
                    // blocks r0 >> SH .. (r0 + 63) >> SH.
                    unsigned long long rows_m = 0;
                    for (int rb = r0 >> SH; rb <= (r0 + 63) >> SH; ++rb) rows_m |= row_mask(rmask, rb);
                    unsigned long long km = row_mask(rmask, j0 >> SH) & rows_m;
                    // Column blocks holding a column left of the panel.
                    const int nbl = (j0 + (1 << SH) - 1) >> SH;
                    if (nbl < 64) km &= (1ull << nbl) - 1;
                    // Groups of sixteen columns go by 1 << GS per block, the
                    // first j0 / 16 of them left of the panel.
                    constexpr int GS = SH - 4;
                    const int ngl = j0 >> 4;
                    const int lr = lane & 15, lk = lane >> 4;
                    int ar[4];
#pragma unroll
                    for (int T = 0; T < 4; ++T) ar[T] = r0 + 16 * T + lr < n ? r0 + 16 * T + lr : n - 1;
                    const int jb = j0 + lr < n ? j0 + lr : n - 1;
                    const bool bon = lr < pw;
                    double ca[4][4], cb[4], na[4][4], nb[4];
                    auto load_group = [&](int k0, double (&a)[4][4], double (&b)[4]) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const auto col = Lt + (size_t)(k0 + 4 * u + lk) * n;
                            b[u] = col[jb];
#pragma unroll
                            for (int T = 0; T < 4; ++T) a[u][T] = col[ar[T]];
                        }
                    };
                    int kb = km ? __builtin_ctzll(km) << GS : 0;
                    bool more = km != 0;
                    if (more) load_group(16 * kb, ca, cb);
                    while (more) {
                        // The next visited group: the next one of the same
                        // block while it is left of the panel, else the first
                        // of the next block (the same group again at the end:
                        // every load stays in bounds and unconditional).
                        int nk = kb + 1;
                        if ((nk & ((1 << GS) - 1)) == 0 || nk >= ngl) {
                            km &= km - 1;
                            more = km != 0;
                            nk = more ? __builtin_ctzll(km) << GS : kb;
                        }
                        load_group(16 * nk, na, nb);
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const double bu = bon ? cb[u] : 0.0;
#pragma unroll
                            for (int T = 0; T < 4; ++T) t[T] = __builtin_amdgcn_mfma_f64_16x16x4f64(bu, ca[u][T], t[T], 0, 0, 0);
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            cb[u] = nb[u];
#pragma unroll
                            for (int T = 0; T < 4; ++T) ca[u][T] = na[u][T];
                        }
                        kb = nk;
                    }
                    // To one row per lane: lane 16 q + m holds rows 16 T + m
                    // (all T) at columns 4 r + q; lane 16 q + m wants row
                    // 16 q + m at every column.  Round s swaps with lane
                    // l ^ 16 s: each lane sends tile q ^ s and receives row
                    // 16 q + m at columns 4 r + (q ^ s).
                    const int q = lane >> 4;
                    double rv[4][4];
#pragma unroll
                    for (int sx = 0; sx < 4; ++sx) {
                        const int tq = q ^ sx;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const double v = tq == 0 ? t[0][r] : tq == 1 ? t[1][r] : tq == 2 ? t[2][r] : t[3][r];
                            rv[sx][r] = sx == 0 ? v : __shfl_xor(v, 16 * sx, 64);
                        }
                    }
#pragma unroll
                    for (int c = 0; c < P; ++c) {
                        const int sx = (c & 3) ^ q, r = c >> 2;
                        const double v = sx == 0 ? rv[0][r] : sx == 1 ? rv[1][r] : sx == 2 ? rv[2][r] : rv[3][r];
                        acc[c] -= v;
                    }
                    k = j0;
                }
                for (; k + 4 <= j0; k += 4) {
                    double li[4], lj[4][P];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        li[u] = Lt[(size_t)(k + u) * n + ii];
#pragma unroll
                        for (int c = 0; c < P; ++c) lj[u][c] = c < pw ? Lt[(size_t)(k + u) * n + j0 + c] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int c = 0; c < P; ++c) acc[c] -= li[u] * lj[u][c];
                }
                for (; k < j0; ++k) {
                    const double li = Lt[(size_t)k * n + ii];
#pragma unroll
                    for (int c = 0; c < P; ++c) acc[c] -= li * (c < pw ? Lt[(size_t)k * n + j0 + c] : 0.0);
                }
                // In-panel columns j = j0 + c, in order.  Rows of the first
                // chunk include the panel's own rows, whose entries are the
                // pivots' and the later columns' multipliers.  A row's
                // diagonal sum stays in a register while the panel runs.
                double q[P];
#pragma unroll
                for (int c = 0; c < P; ++c) {
                    q[c] = 0.0;
                    if (c < pw) {
                        const int j = j0 + c;
                        double sum = acc[c];
#pragma unroll
                        for (int c2 = 0; c2 < P; ++c2)
                            if (c2 < c) {
                                // L[j][j0 + c2]: computed earlier in this panel (row j is
                                // in the first chunk: its lane's q[c2], else from Lt).
                                const double ljk = readlane_f64(r0 == j0 ? q[c2] : q0[c2], j - j0);
                                sum -= q[c2] * ljk;
                            }
                        // Pivot L[j][j] = sqrt(pt[j]) once every k < j is subtracted.
                        double ljj;
                        if (r0 == j0) {
                            const double ptj = readlane_f64(pr, j - j0);
                            if (ptj <= 0.0) return 4;  // NEWT_ERR_NEWTON_FAIL: Hessian not positive definite
                            ljj = __builtin_sqrt(ptj);
                        } else {
                            ljj = readlane_f64(q0[c], j - j0);
                        }
                        if (on && i > j) {
                            q[c] = sum / ljj;
                            pr -= q[c] * q[c];
                        } else if (on && i == j) {
                            q[c] = ljj;
                        }
                    }
                }
                if (mfma_left) {
                    bool nz = false;
#pragma unroll
                    for (int c = 0; c < P; ++c) nz |= q[c] != 0.0;
                    const unsigned long long bm = __ballot(nz);
                    // Rows r0 + 16 g .. r0 + 16 g + 15 lie in block (r0 + 16 g) >> SH.
#pragma unroll
                    for (int g = 0; g < 4; ++g)
                        if (lane == (r0 + 16 * g) >> SH && ((bm >> (16 * g)) & 0xffffull) != 0)
                            rmask |= 1ull << (j0 >> SH);
                }
#pragma unroll
                for (int c = 0; c < P; ++c) {
                    const int j = j0 + c;
                    if (c < pw && on && i >= j) {
                        Lt[(size_t)j * n + i] = q[c];
                        Lf[(size_t)i * n + j] = q[c];
                    }
                }
                if (r0 == j0) {
#pragma unroll
                    for (int c = 0; c < P; ++c) q0[c] = q[c];
                }
                if (on) pt[i] = pr;
                // The panel's chunks read nothing their siblings write (other
                // rows' pt, columns left of the panel): one fence per panel.
                if (r0 + 64 >= n) gsync();
            }
        }
        const uint32_t max_it = m.iterations < 1 ? 1u : (uint32_t)m.iterations;
        // Probe points P(x + a d) and their products with H, four per pass
        // over H; hcur holds H x for the current x (the gradient's product,
        // bitwise what mat_vec(H, x) gives), hbest H times the best probe.
        gf64 *vb = V + 10 * nn, *ob = V + 14 * nn, *hcur = hv, *hbest = V + 18 * nn;
        double cur = cost(H, lin, x, hv, n);
        for (uint32_t it = 0; it < max_it; ++it) {
            // gradient, direction = cholesky_solve(L, -g) (dynamics.rs:626-648)
            for (int i = lane; i < n; i += 64) {
                g[i] = hcur[i] + lin[i];
                y[i] = -g[i];
            }
            gsync();
            // forward: y[i] -= L[i][k] y[k] for k = 0, 1, ... (the row sum's
            // order), y[64 c + lane] held in register yr[c], sixteen columns
            // k0 .. k0+15 at a time: the panel's own rows (all in one chunk)
            // are solved in order with the pivot's owner dividing and its
            // value broadcast, then every row below takes the panel's sixteen
            // terms in k order, all sixteen factor loads of a row in flight
            // together.
            {
                double yr[C];
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int i = 64 * c + lane;
                    yr[c] = i < n ? y[i] : 0.0;
                }
                for (int k0 = 0; k0 < n; k0 += 16) {
                    const int kw = n - k0 < 16 ? n - k0 : 16, cp = k0 >> 6, base = k0 & 63;
                    const int rp = 64 * cp + lane;  // this lane's row in the panel's chunk
                    const int rpc = rp < n ? rp : n - 1;
                    double lp[16], yp[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) lp[u] = u < kw ? Lt[(size_t)(k0 + u) * n + rpc] : 0.0;
                    double ycp = 0.0;
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        if (c == cp) ycp = yr[c];
#pragma unroll
                    for (int u = 0; u < 16; ++u) {
                        yp[u] = 0.0;
                        if (u < kw) {
                            const int k = k0 + u, own = base + u;
                            const double yk = readlane_f64(ycp / lp[u], own);
                            if (lane == own) ycp = yk;
                            else if (rp > k && rp < n) ycp -= lp[u] * yk;
                            yp[u] = yk;
                        }
                    }
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        if (c == cp) yr[c] = ycp;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int i = 64 * c + lane;
                        if (c > cp && 64 * c < n) {
                            const int ic = i < n ? i : n - 1;
                            double l[16];
#pragma unroll
                            for (int u = 0; u < 16; ++u) l[u] = u < kw ? Lt[(size_t)(k0 + u) * n + ic] : 0.0;
                            if (i < n) {
#pragma unroll
                                for (int u = 0; u < 16; ++u)
                                    if (u < kw) yr[c] -= l[u] * yp[u];
                            }
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int i = 64 * c + lane;
                    if (i < n) y[i] = yr[c];
                }
            }
            gsync();
            if (!exact) {
                // backward, sixteen rows i0 .. i0+15 at a time from the bottom:
                // x[i] = y[i] / L[i][i] for the panel's rows in decreasing i,
                // each then subtracted from the panel rows above it, then
                // y[k] -= L[i][k] x[i] for every k < i0, the sixteen factor
                // rows' loads in flight together.  Each sum takes its terms in
                // decreasing i, not newt's increasing order: within the
                // drop-in's tolerance.
                double yr[C];
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int k = 64 * c + lane;
                    yr[c] = k < n ? y[k] : 0.0;
                }
                for (int i0 = (n - 1) & ~15; i0 >= 0; i0 -= 16) {
                    const int iw = n - i0 < 16 ? n - i0 : 16, cp = i0 >> 6, base = i0 & 63;
                    const int kp = 64 * cp + lane;  // this lane's index in the panel's chunk
                    const int kpc = kp < n ? kp : n - 1;
                    double lp[16], xp[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) lp[u] = u < iw ? Lf[(size_t)(i0 + u) * n + kpc] : 0.0;
                    double ycp = 0.0;
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        if (c == cp) ycp = yr[c];
#pragma unroll
                    for (int u = 15; u >= 0; --u) {
                        xp[u] = 0.0;
                        if (u < iw) {
                            const int i = i0 + u, own = base + u;
                            const double xi = readlane_f64(ycp / lp[u], own);
                            if (lane == own) ycp = xi;
                            else if (kp < i) ycp -= lp[u] * xi;
                            xp[u] = xi;
                        }
                    }
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        if (c == cp) yr[c] = ycp;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int k = 64 * c + lane;
                        if (c < cp) {
                            double l[16];
#pragma unroll
                            for (int u = 0; u < 16; ++u) l[u] = u < iw ? Lf[(size_t)(i0 + u) * n + k] : 0.0;
#pragma unroll
                            for (int u = 15; u >= 0; --u)
                                if (u < iw) yr[c] -= l[u] * xp[u];
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int k = 64 * c + lane;
                    if (k < n) dd[k] = yr[c];
                }
            } else
            // backward: x[i] = (y[i] - sum_{k>i} L[k][i] x[k]) / L[i][i], the sum
            // in increasing k.  Lane k mod 64 keeps x[k] in a register (chunk
            // k / 64 of at most C) and forms the product L[k][i] x[k]; every
            // lane runs the subtraction chain on the broadcast products in k
            // order.  Row i - 1's factor entries, y[i - 1] and L[i-1][i-1] load
            // while row i is summed, so no load sits on the chain.
            {
                double xr[C], lr[C], ln[C];
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    xr[c] = 0.0;
                    lr[c] = 0.0;
                }
                auto load_row = [&](int i, double* out) {
                    const auto lti = Lt + (size_t)i * n;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int k = 64 * c + lane;
                        out[c] = (k > i && k < n) ? lti[k] : 0.0;
                    }
                };
                double yi = 0.0, di = 1.0;
                if (n > 0) {
                    yi = y[n - 1];
                    di = Lt[(size_t)(n - 1) * n + (n - 1)];
                    load_row(n - 1, lr);
                }
                for (int i = n - 1; i >= 0; --i) {
                    double ny = 0.0, nd = 1.0;
                    if (i > 0) {
                        ny = y[i - 1];
                        nd = Lt[(size_t)(i - 1) * n + (i - 1)];
                        load_row(i - 1, ln);
                    }
                    double sum = yi;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int b0 = 64 * c;
                        if (b0 + 63 <= i || b0 >= n) continue;
                        const double p = lr[c] * xr[c];
                        const int lo = (i + 1 > b0 ? i + 1 : b0) - b0;
                        const int hi = (n < b0 + 64 ? n : b0 + 64) - b0;
                        if (lo == 0 && hi == 64) {
#pragma unroll
                            for (int u = 0; u < 64; ++u) sum -= readlane_f64(p, u);
                        } else {
#pragma unroll
                            for (int u = 0; u < 64; ++u) {
                                const double pu = readlane_f64(p, u);
                                if (u >= lo && u < hi) sum -= pu;
                            }
                        }
                    }
                    const double xi = sum / di;
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        if (64 * c + lane == i) xr[c] = xi;
                    if (lane == (i & 63)) dd[i] = xi;
                    if (i > 0) {
                        yi = ny;
                        di = nd;
#pragma unroll
                        for (int c = 0; c < C; ++c) lr[c] = ln[c];
                    }
                }
            }
            gsync();
            if (dot_seq(g, dd, n) >= 0.0) break;
            // exact_pyramidal_line_search with NonNegative projections
            // (newton.rs:400-556).  Breakpoints: 0, 1, then -x[i] / d[i] in
            // (0, 1) in index order (compacted across the wave), sorted and
            // deduplicated on lane 0.
            int nbk = 0;
            {
                int c = 2;
                for (int i0 = 0; i0 < n; i0 += 64) {
                    const int i = i0 + lane;
                    double alpha = 0.0;
                    bool ok = false;
                    if (i < n) {
                        const double slope = dd[i];
                        if (slope != 0.0) {
                            alpha = -x[i] / slope;
                            ok = alpha > 0.0 && alpha < 1.0;
                        }
                    }
                    const unsigned long long mk = __ballot(ok);
                    if (ok) brk[c + __popcll(mk & lanes_below(lane))] = alpha;
                    c += __popcll(mk);
                }
                if (lane == 0) {
                    brk[0] = 0.0;
                    brk[1] = 1.0;
                }
                gsync();
                if (lane == 0) {
                    for (int a = 1; a < c; ++a) {  // stable insertion sort
                        const double v = brk[a];
                        int b = a - 1;
                        while (b >= 0 && brk[b] > v) {
                            brk[b + 1] = brk[b];
                            --b;
                        }
                        brk[b + 1] = v;
                    }
                    int u = 0;
                    for (int a = 0; a < c; ++a) {
                        const double v = brk[a];
                        if (u > 0 && __builtin_fabs(v - brk[u - 1]) <= 1e-7) continue;
                        brk[u++] = v;
                    }
                    nbk = u;
                }
                nbk = __shfl(nbk, 0, 64);
                gsync();
            }
            int ncand = nbk;
            for (int a = lane; a < nbk; a += 64) cand[a] = brk[a];
            gsync();
            // Each interval's stationary point from the slope of the projected
            // path (mid point and secant slope), two intervals per pass over H.
            for (int k = 0; k + 1 < nbk;) {
                int nk = 0;
                double mids[2], los[2], his[2];
                while (k + 1 < nbk && nk < 2) {
                    const double lo = brk[k], hi = brk[k + 1];
                    ++k;
                    if (hi - lo <= 1e-7) continue;
                    const double mid = 0.5 * (lo + hi);
                    const double eps = (hi - lo) * 1e-4;
                    const double la = nmax(mid - eps, lo);
                    const double ra = nmin(mid + eps, hi);
                    const double span = ra - la;
                    if (span <= 0.0) continue;
                    const double inv = 1.0 / span;
                    for (int i = lane; i < n; i += 64) {
                        const double xi = x[i], di = dd[i];
                        const double vl = xi + la * di, vr = xi + ra * di, vm = xi + mid * di;
                        const double pli = vl < 0.0 ? 0.0 : vl, pri = vr < 0.0 ? 0.0 : vr;
                        vb[(size_t)(2 * nk) * nn + i] = vm < 0.0 ? 0.0 : vm;
                        vb[(size_t)(2 * nk + 1) * nn + i] = (pri - pli) * inv;
                    }
                    mids[nk] = mid;
                    los[nk] = lo;
                    his[nk] = hi;
                    ++nk;
                }
                if (nk == 0) continue;
                gsync();
                mat_vec4(H, vb, ob, 2 * nk, n, nn);
                for (int t = 0; t < nk; ++t) {
                    gf64 *pm = vb + (size_t)(2 * t) * nn, *sl_t = vb + (size_t)(2 * t + 1) * nn;
                    gf64 *hm = ob + (size_t)(2 * t) * nn, *hs = ob + (size_t)(2 * t + 1) * nn;
                    for (int i = lane; i < n; i += 64) pm[i] = hm[i] + lin[i];  // local gradient
                    gsync();
                    const double den = dot_seq(sl_t, hs, n);
                    const double num = dot_seq(pm, sl_t, n);
                    const double mid = mids[t], lo = los[t], hi = his[t];
                    const double opt = den > 0.0 ? nmin(nmax(mid - num / den, lo), hi) : mid;
                    if (lane == 0) cand[ncand] = opt;
                    ++ncand;
                }
                gsync();
            }
            // Every candidate's cost 0.5 pᵀHp + linᵀp in candidate order (the
            // first strict minimum wins), four probes per pass over H.
            double best_a = 0.0, best_c = cur;
            for (int c0 = 0; c0 < ncand; c0 += 4) {
                const int K = ncand - c0 < 4 ? ncand - c0 : 4;
                for (int t = 0; t < K; ++t) affine(x, dd, cand[c0 + t], vb + (size_t)t * nn, n);
                mat_vec4(H, vb, ob, K, n, nn);
                for (int t = 0; t < K; ++t) {
                    gf64 *pt_t = vb + (size_t)t * nn, *hp = ob + (size_t)t * nn;
                    const double cc = 0.5 * dot_seq(pt_t, hp, n) + dot_seq(lin, pt_t, n);
                    if (cc < best_c) {
                        best_c = cc;
                        best_a = cand[c0 + t];
                        for (int i = lane; i < n; i += 64) hbest[i] = hp[i];
                    }
                }
                gsync();
            }
            if (best_a <= 0.0 || cur - best_c <= 1e-7 * (1.0 + __builtin_fabs(cur))) break;
            for (int i = lane; i < n; i += 64) {
                const double v = x[i] + best_a * dd[i];
                x[i] = v < 0.0 ? 0.0 : v;
            }
            {
                gf64* t = hcur;
                hcur = hbest;
                hbest = t;
            }
            gsync();
            cur = best_c;
        }
        for (int i = lane; i < n; i += 64) rows[7 * i + 6] = x[i];
        gsync();
        return 0;
    }

    // solve_free_bodies (solver.rs:703-1257): wrenches into LDS; 0 or a
    // fault code.
    __device__ __forceinline__ int solve(int nc) const {
        const int nb = m.n_bodies;
        for (int i = lane; i < 6 * nb; i += 64) sd[B.wr + i] = 0.0;
        if (nc == 0 || m.dt <= 0.0) {
            gsync();
            return 0;
        }
        for (int i = lane; i < nb; i += 64) {
            M33<double> r = qmat(bori(i));
            M33<double> w3 = FreeBodies::mm(FreeBodies::mm(r, FreeBodies::tbl(body(i).inv_inertia)), FreeBodies::tr(r));
            for (int k = 0; k < 9; ++k) sd[B.iiw + 9 * i + k] = w3.m[k];
            for (int k = 0; k < 6; ++k) sd[B.bd + 6 * i + k] = 0.0;
        }
        // Active contacts (a body side, penetration beyond the gap), in order.
        int nblk = 0;
        for (int base = 0; base < nc; base += 64) {
            const int c = base + lane;
            bool act = false;
            if (c < nc) {
                const int a = m.geoms[cid(c, 0)].body, b = m.geoms[cid(c, 1)].body;
                act = (a >= 0 || b >= 0) && cg(c, 6) - cg(c, 8) > 0.0;
            }
            nblk += __popcll(__ballot(act));
        }
        if (nblk <= B.lds_contacts) {
            rows = sd + B.rows;
            blk = sd + B.blk;
            bbi = si + B.bbi;
        } else {
            rows = gd + B.g_rows;
            blk = gd + B.g_blk;
            bbi = gi + B.g_bbi;
        }
        gsync();
        int k0 = 0;
        for (int base = 0; base < nc; base += 64) {
            const int c = base + lane;
            bool act = false;
            if (c < nc) {
                const int a = m.geoms[cid(c, 0)].body, b = m.geoms[cid(c, 1)].body;
                act = (a >= 0 || b >= 0) && cg(c, 6) - cg(c, 8) > 0.0;
            }
            const unsigned long long mask = __ballot(act);
            if (act) assemble(k0 + __popcll(mask & lanes_below(lane)), c);
            k0 += __popcll(mask);
        }
        gsync();
        if (nblk == 0) return 0;
        if (m.mode == NEWT_SOLVER_NEWTON) {
            const int st = nblk <= 256 ? newton<16>(nblk) : newton<32>(nblk);
            if (st != 0) return st;
        } else {
            if (lane == 0)
                body_pgs_sweeps(m.iterations, nblk, rows, blk, bbi, sd + B.bd, sd + B.iiw, sd + B.st, m.bodies);
            gsync();
        }
        // Wrenches, one lane per body, rows in order (solver.rs:1198-1235).
        // Rows go in chunks of 64: lane u forms row r0 + u's force and its
        // torques about both contact points, then every body lane takes the
        // chunk's rows in order from broadcasts and adds those on its body.
        const double dt = m.dt;
        const int nr = 4 * nblk;
        for (int b00 = 0; b00 < nb; b00 += 64) {
            const int b0 = b00 + lane;
            V3<double> f = z3<double>(), t = z3<double>();
            for (int r0 = 0; r0 < nr; r0 += 64) {
                const int ri = r0 + lane < nr ? r0 + lane : nr - 1;
                const int k = ri >> 2;
                const double fi = rows[7 * ri + 6];
                const bool live = r0 + lane < nr && fi != 0.0;
                const int ra = live ? bbi[2 * k] : -2, rb = live ? bbi[2 * k + 1] : -2;  // -2: no body
                const V3<double> fa = rdir(ri) * (fi / dt), fb = -fa;
                const V3<double> ta = cross3(ld3(blk + 8 * k), fa), tb = cross3(ld3(blk + 8 * k + 3), fb);
                const int cnt = nr - r0 < 64 ? nr - r0 : 64;
                for (int u = 0; u < cnt; ++u) {
                    const int au = __builtin_amdgcn_readlane(ra, u), bu = __builtin_amdgcn_readlane(rb, u);
                    if (au == -2) continue;  // uniform
                    const V3<double> fau = mk3(readlane_f64(fa.x, u), readlane_f64(fa.y, u), readlane_f64(fa.z, u));
                    const V3<double> tau = mk3(readlane_f64(ta.x, u), readlane_f64(ta.y, u), readlane_f64(ta.z, u));
                    const V3<double> fbu = mk3(readlane_f64(fb.x, u), readlane_f64(fb.y, u), readlane_f64(fb.z, u));
                    const V3<double> tbu = mk3(readlane_f64(tb.x, u), readlane_f64(tb.y, u), readlane_f64(tb.z, u));
                    if (au == b0) {
                        f = f + fau;
                        t = t + tau;
                    }
                    if (bu == b0) {
                        f = f + fbu;
                        t = t + tbu;
                    }
                }
            }
            if (b0 < nb) {
                st3(sd + B.wr + 6 * b0, f);
                st3(sd + B.wr + 6 * b0 + 3, t);
            }
        }
        gsync();
        return 0;
    }

    // ---------------- integration (world.rs:1705-1817, 3223-3286) ---------
    __device__ __forceinline__ FreeBodies::Dv evaluate(int b, const FreeBodies::St& s, V3<double> f, V3<double> tq) const {
        const BodyT& bt = body(b);
        FreeBodies::Dv dv;
        dv.dv = gravity_of(b) + f / bt.mass;
        const V3<double> iwv = mv(FreeBodies::tbl(bt.inertia), s.om);
        const V3<double> gyro = -cross3(s.om, iwv);
        const V3<double> tb = qrot_inv(s.q, tq);
        dv.dw = mv(FreeBodies::tbl(bt.inv_inertia), tb + gyro);
        dv.dq = qderiv(s.q, s.om);
        dv.dp = s.v;
        return dv;
    }
    __device__ __forceinline__ void integrate() const {
        using FB = FreeBodies;
        const double dt = m.dt;
        const size_t W = d.W;
        for (int b = lane; b < m.n_bodies; b += 64) {
            const BodyT& bt = body(b);
            const FB::St s0{bpos(b), bvel(b), bori(b), bomega(b)};
            const V3<double> f = ld3(sd + B.wr + 6 * b), tq = ld3(sd + B.wr + 6 * b + 3);
            const FB::Dv k1 = evaluate(b, s0, f, tq);
            FB::St n;
            if (m.integrator == NEWT_INTEGRATOR_RK4) {
                const FB::Dv k2 = evaluate(b, FB::advance(s0, k1, dt * 0.5), f, tq);
                const FB::Dv k3 = evaluate(b, FB::advance(s0, k2, dt * 0.5), f, tq);
                const FB::Dv k4 = evaluate(b, FB::advance(s0, k3, dt), f, tq);
                const double sixth = 1.0 / 6.0;
                const V3<double> dp = (((k1.dp + k2.dp * 2.0) + k3.dp * 2.0) + k4.dp) * sixth;
                const V3<double> dv = (((k1.dv + k2.dv * 2.0) + k3.dv * 2.0) + k4.dv) * sixth;
                const Qt<double> dq =
                    FB::qsc(FB::qadd(FB::qadd(FB::qadd(k1.dq, FB::qsc(k2.dq, 2.0)), FB::qsc(k3.dq, 2.0)), k4.dq), sixth);
                const V3<double> dwv = (((k1.dw + k2.dw * 2.0) + k3.dw * 2.0) + k4.dw) * sixth;
                n.p = s0.p + dp * dt;
                n.v = FB::clampv(s0.v + dv * dt, bt.has_max_lin, bt.max_lin);
                n.q = qnormalize(FB::qadd(s0.q, FB::qsc(dq, dt)));
                n.om = FB::clampv(s0.om + dwv * dt, bt.has_max_ang, bt.max_ang);
            } else {
                const V3<double> v = s0.v + k1.dv * dt;
                const V3<double> om = s0.om + k1.dw * dt;
                n.p = s0.p + v * dt;
                n.q = qintegrate(s0.q, om, dt);
                n.v = FB::clampv(v, bt.has_max_lin, bt.max_lin);
                n.om = FB::clampv(om, bt.has_max_ang, bt.max_ang);
            }
            double* s = sd + B.st + 13 * b;
            st3(s, n.p);
            s[3] = n.q.x; s[4] = n.q.y; s[5] = n.q.z; s[6] = n.q.w;
            st3(s + 7, n.v);
            st3(s + 10, n.om);
            const int vo = bt.v_off;
            d.qacc[(size_t)vo * W + world] = k1.dv.x;
            d.qacc[(size_t)(vo + 1) * W + world] = k1.dv.y;
            d.qacc[(size_t)(vo + 2) * W + world] = k1.dv.z;
            d.qacc[(size_t)(vo + 3) * W + world] = k1.dw.x;
            d.qacc[(size_t)(vo + 4) * W + world] = k1.dw.y;
            d.qacc[(size_t)(vo + 5) * W + world] = k1.dw.z;
        }
        gsync();
    }

    // World::step; 0 or a newt_status fault code.  np / nc: pair and
    // contact counts of the step.
    __device__ __forceinline__ int step(int* np, int* nc) const {
        *nc = 0;
        *np = 0;
        int c = 0;
        if (m.mode == NEWT_SOLVER_PGS || m.mode == NEWT_SOLVER_NEWTON) {
            c = contacts(np);
            if (c > m.cap_c) return 8;  // NEWT_ERR_CAPACITY
        } else {
            // active pairs only (penalty models here have no candidate pairs)
            *np = 0;
        }
        *nc = c;
        const int st = solve(c);
        if (st != 0) return st;
        integrate();
        populate(false);
        return 0;
    }
};

}  // namespace newt
