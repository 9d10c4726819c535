// newt_sensors.hpp — sensor::evaluate (newt/src/sensor.rs:446-702) on the
// post-step state, one lane per world, FP64.
//
// newt evaluates its sensors after every step and keeps the last values
// (world.rs:1516-1522); evaluation only reads the state, so running it once
// after the last step of a launch gives the same sensordata.  Tree models
// use the lane step's workspace (World<R>): forward kinematics, and for the
// acceleration and force sensors the start-of-step contacts of the last
// step re-solved at the post-step state (World::build_sensor_inputs,
// world.rs:1578-1696) followed by an explicit ABA with the contact wrenches.
// Free-body models read their bodies' state directly.
#pragma once

#include "newt_step.hpp"

namespace newt {

template <class R> struct SensorEval {
    World<R>& wd;
    const ModelT& m;
    const Layout& L;
    Lane<double> out;  // this world's sensordata, [sensor_dim]

    NEWT_HD V3<double> ld3(int o) const { return {wd.dw[o], wd.dw[o + 1], wd.dw[o + 2]}; }
    NEWT_HD void st3(int o, V3<double> v) const { wd.dw[o] = v.x; wd.dw[o + 1] = v.y; wd.dw[o + 2] = v.z; }
    NEWT_HD SM<double> ldm(int o) const { return {ld3(o), ld3(o + 3)}; }
    NEWT_HD void stm(int o, SM<double> v) const { st3(o, v.w); st3(o + 3, v.v); }
    NEWT_HD SF<double> ldf(int o) const { return {ld3(o), ld3(o + 3)}; }
    NEWT_HD void stf(int o, SF<double> v) const { st3(o, v.t); st3(o + 3, v.f); }
    NEWT_HD void put(int o, V3<double> v) const { out[o] = v.x; out[o + 1] = v.y; out[o + 2] = v.z; }

    // compute_link_va (sensor.rs:834-916): per link (v, a) in its body frame
    // at the COM into dw[L.sva + 12 i]; qddot from dw[qdd], or zero when
    // qdd < 0 (only v is read then).
    NEWT_HD void link_va(int qdd) const {
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            SM<double> v, a;
            if (l.joint == JF_FREE) {
                const int o = l.v_off;
                v = {mk3(wd.qvel[o], wd.qvel[o + 1], wd.qvel[o + 2]), mk3(wd.qvel[o + 3], wd.qvel[o + 4], wd.qvel[o + 5])};
                a = qdd < 0 ? smz<double>()
                            : SM<double>{mk3(wd.dw[qdd + o], wd.dw[qdd + o + 1], wd.dw[qdd + o + 2]),
                                         mk3(wd.dw[qdd + o + 3], wd.dw[qdd + o + 4], wd.dw[qdd + o + 5])};
            } else {
                const Xf<double> x = wd.xup(i, false);
                const SM<double> vp = l.parent >= 0 ? ldm(L.sva + 12 * l.parent) : smz<double>();
                const SM<double> ap = l.parent >= 0 ? ldm(L.sva + 12 * l.parent + 6) : smz<double>();
                if (l.joint == JF_FIXED) {
                    v = xmotion(x, vp);
                    a = xmotion(x, ap);
                } else {
                    // propagate_single_dof (sensor.rs:935-962)
                    const SM<double> s = wd.subspace(i);
                    const SM<double> sq = s * wd.qvel[l.v_off];
                    const SM<double> sqq = s * (qdd < 0 ? 0.0 : wd.dw[qdd + l.v_off]);
                    v = xmotion(x, vp) + sq;
                    const SM<double> c = crm(v, sq);
                    a = (xmotion(x, ap) + sqq) + c;
                }
            }
            stm(L.sva + 12 * i, v);
            stm(L.sva + 12 * i + 6, a);
        }
    }

    // The re-solve of the last step's contacts at the post-step state
    // (solve_tree_contacts via solve_tree_contact_sensor_solution,
    // world.rs:2252-2274) and its per-link world wrenches and per-contact
    // normal forces (solver.rs:3334-3364): rows in contact order, sides A
    // then B, wrench (dir·sign, (p − com) × dir·sign) per unit force; the
    // normal force of a pyramidal block is its four impulses summed / dt.
    // False when the post-step mass matrix is not positive definite.
    NEWT_HD bool contact_forces(int nc) const {
        for (int k = 0; k < 6 * m.n_links; ++k) wd.dw[L.swr + k] = 0.0;
        for (int k = 0; k < nc; ++k) wd.dw[L.sfn + k] = 0.0;
        if (nc == 0 || m.mode != 1) return true;  // PGS only (the other modes have no pairs here)
        int nr = 0;
        if (wd.solve_contacts(nc, (R)m.dt, &nr) != 0) return false;
        const double dt = m.dt;
        int row = 0;
        for (int k = 0; k < nc; ++k) {
            const int o = L.cgeo + 9 * k;
            const double pen = wd.dw[o + 6] - wd.dw[o + 8];
            if (pen <= 0.0) continue;
            const GeomT& ga = m.geoms[wd.iw[L.cids + 3 * k]];
            const GeomT& gb = m.geoms[wd.iw[L.cids + 3 * k + 1]];
            const V3<double> cp = ld3(o), n = ld3(o + 3);
            const double mu = ga.friction < gb.friction ? ga.friction : gb.friction;
            const V3<double> ref = nabs(n.z) < 0.9 ? mk3(0.0, 0.0, 1.0) : mk3(1.0, 0.0, 0.0);
            const V3<double> t1 = unit3(cross3(ref, n));
            const V3<double> t2 = cross3(n, t1);
            const V3<double> dirs[4] = {n + t1 * mu, n - t1 * mu, n + t2 * mu, n - t2 * mu};
            double fsum = -0.0;
            for (int r = 0; r < 4; ++r) {
                const double fr = wd.dw[L.f + row + r];
                fsum = fsum + fr;
                const double force = fr / dt;
                for (int side = 0; side < 2; ++side) {
                    const GeomT& g = side == 0 ? ga : gb;
                    if (g.link < 0) continue;
                    const V3<double> d = dirs[r] * (side == 0 ? 1.0 : -1.0);
                    const int wo = L.swr + 6 * g.link;
                    st3(wo, ld3(wo) + d * force);
                    st3(wo + 3, ld3(wo + 3) + cross3(cp - wd.pos(g.link), d) * force);
                }
            }
            wd.dw[L.sfn + k] = fsum / dt;
            row += 4;
        }
        return true;
    }

    // compute_link_wrenches (sensor.rs:964-1028): per link the wrench the
    // parent joint transmits, child frame at the COM, into dw[L.sfl + 6 i].
    NEWT_HD void link_wrenches() const {
        const V3<double> g = mk3(m.gravity[0], m.gravity[1], m.gravity[2]);
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            const SM<double> v = ldm(L.sva + 12 * i), a = ldm(L.sva + 12 * i + 6);
            const SF<double> ia = m6mul(l.si, a), iv = m6mul(l.si, v);
            const SF<double> bias = crf(v, iv);
            const V3<double> fw = ld3(L.swr + 6 * i) + g * l.mass;
            const Qt<double> o = wd.ori(i);
            const SF<double> fe{qrot_inv(o, ld3(L.swr + 6 * i + 3)), qrot_inv(o, fw)};
            stf(L.sfl + 6 * i, (ia + bias) - fe);
        }
        for (int i = m.n_links - 1; i >= 1; --i) {
            const int p = m.links[i].parent;
            const Xf<double> x = wd.xup(i, false);
            stf(L.sfl + 6 * p, ldf(L.sfl + 6 * p) + xtforce(x, ldf(L.sfl + 6 * i)));
        }
    }

    // site_world_pose (sensor.rs:794-808) of a site on a link or a free body.
    NEWT_HD void parent_pose(const SensorT& s, V3<double>* p, Qt<double>* q) const {
        if (s.link >= 0) {
            *p = wd.pos(s.link);
            *q = wd.ori(s.link);
        } else {
            const BodyT& b = m.bodies[s.body];
            const int o = b.q_off;
            *p = mk3(wd.qpos[o], wd.qpos[o + 1], wd.qpos[o + 2]);
            *q = Qt<double>{wd.qpos[o + 3], wd.qpos[o + 4], wd.qpos[o + 5], wd.qpos[o + 6]};
        }
    }

    // (ω_body, v_body) of the site's parent (parent_body_velocities,
    // sensor.rs:814-832).
    NEWT_HD void parent_vel(const SensorT& s, V3<double>* w, V3<double>* v) const {
        if (s.link >= 0) {
            *w = ld3(L.sva + 12 * s.link);
            *v = ld3(L.sva + 12 * s.link + 3);
        } else {
            const BodyT& b = m.bodies[s.body];
            const int o = b.v_off, q = b.q_off;
            const Qt<double> ori{wd.qpos[q + 3], wd.qpos[q + 4], wd.qpos[q + 5], wd.qpos[q + 6]};
            *w = mk3(wd.qvel[o + 3], wd.qvel[o + 4], wd.qvel[o + 5]);
            *v = qrot_inv(ori, mk3(wd.qvel[o], wd.qvel[o + 1], wd.qvel[o + 2]));
        }
    }

    // site_velocity (sensor.rs:705-726): (linear, angular) in world.
    NEWT_HD void site_velocity(const SensorT& s, V3<double>* lin, V3<double>* ang) const {
        const V3<double> off = mk3(s.off[0], s.off[1], s.off[2]);
        if (s.link >= 0) {
            const V3<double> vw = ld3(L.sva + 12 * s.link), vv = ld3(L.sva + 12 * s.link + 3);
            const Qt<double> o = wd.ori(s.link);
            *ang = qrot(o, vw);
            *lin = qrot(o, vv + cross3(vw, off));
        } else {
            const BodyT& b = m.bodies[s.body];
            const int vo = b.v_off, q = b.q_off;
            const Qt<double> o{wd.qpos[q + 3], wd.qpos[q + 4], wd.qpos[q + 5], wd.qpos[q + 6]};
            *ang = qrot(o, mk3(wd.qvel[vo + 3], wd.qvel[vo + 4], wd.qvel[vo + 5]));
            *lin = mk3(wd.qvel[vo], wd.qvel[vo + 1], wd.qvel[vo + 2]) + cross3(*ang, qrot(o, off));
        }
    }

    // Every sensor of the model.  `nc` = the last step's contacts (start of
    // that step), in the workspace's contact list.
    NEWT_HD void evaluate(int nc) const {
        bool va = false, dyn = false;
        for (int k = 0; k < m.n_sensors; ++k) {
            va = va || sensor_kind_va(m.sensors[k].kind);
            dyn = dyn || sensor_kind_dynamic(m.sensors[k].kind);
        }
        const bool tree = m.n_links > 0;  // trees, alone or beside free bodies
        if (tree) wd.forward_kinematics();
        bool ok = true;
        if (tree && dyn) {
            ok = contact_forces(nc);
            if (ok) {
                // aba (tree.rs:942-955): explicit, the user's applied force,
                // the contact wrenches as external forces, penalty limits on
                // (the step has restored disable_penalty_limits).
                for (int s = 0; s < m.tnv; ++s) wd.dw[L.tau + s] = wd.qfrc[s];
                wd.ext_wr = L.swr;
                wd.aba(false, false, m.dt, L.sqdd);
                wd.ext_wr = -1;
            }
        }
        if (tree && va) link_va(dyn && ok ? L.sqdd : -1);
        if (tree && dyn && ok) link_wrenches();
        const V3<double> g = mk3(m.gravity[0], m.gravity[1], m.gravity[2]);
        const double nan = __builtin_nan("");
        for (int k = 0; k < m.n_sensors; ++k) {
            const SensorT& s = m.sensors[k];
            const int o = s.out;
            if (sensor_kind_dynamic(s.kind) && !ok) {
                for (int j = 0; j < sensor_kind_dim(s.kind); ++j) out[o + j] = nan;
                continue;
            }
            const V3<double> off = mk3(s.off[0], s.off[1], s.off[2]);
            const Qt<double> lq{s.quat[0], s.quat[1], s.quat[2], s.quat[3]};
            switch (s.kind) {
                case SK_JOINT_POS: out[o] = wd.qpos[m.links[s.link].q_off]; break;
                case SK_JOINT_VEL: out[o] = wd.qvel[m.links[s.link].v_off]; break;
                case SK_FRAME_POS: {
                    V3<double> p; Qt<double> q;
                    parent_pose(s, &p, &q);
                    put(o, p + qrot(q, off));
                    break;
                }
                case SK_FRAME_QUAT: {
                    V3<double> p; Qt<double> q;
                    parent_pose(s, &p, &q);
                    const Qt<double> r = qmul(q, lq);
                    out[o] = r.x; out[o + 1] = r.y; out[o + 2] = r.z; out[o + 3] = r.w;
                    break;
                }
                case SK_GYRO: {
                    V3<double> w, v;
                    parent_vel(s, &w, &v);
                    put(o, qrot_inv(lq, w));
                    break;
                }
                case SK_ACCEL: {
                    // accelerometer_reading (sensor.rs:1030-1088), link sites
                    const Qt<double> ori = wd.ori(s.link);
                    const SM<double> vb = ldm(L.sva + 12 * s.link), ab = ldm(L.sva + 12 * s.link + 6);
                    const V3<double> om = qrot(ori, vb.w), al = qrot(ori, ab.w);
                    const V3<double> acom = qrot(ori, ab.v + cross3(vb.w, vb.v));
                    const V3<double> r = qrot(ori, off);
                    const V3<double> asite = (acom + cross3(al, r)) + cross3(om, cross3(om, r));
                    put(o, qrot_inv(qmul(ori, lq), asite - g));
                    break;
                }
                case SK_TOUCH: {
                    // touch_reading (sensor.rs:1103-1112)
                    double sum = 0.0;
                    for (int c = 0; c < nc; ++c)
                        if (wd.iw[L.cids + 3 * c] == s.geom || wd.iw[L.cids + 3 * c + 1] == s.geom)
                            sum = sum + nabs(wd.dw[L.sfn + c]);
                    out[o] = sum;
                    break;
                }
                case SK_FORCE: put(o, ld3(L.sfl + 6 * s.link + 3)); break;
                case SK_TORQUE: {
                    const SF<double> f = ldf(L.sfl + 6 * s.link);
                    const V3<double> rj = mk3(m.links[s.link].cj[0], m.links[s.link].cj[1], m.links[s.link].cj[2]);
                    put(o, f.t - cross3(rj, f.f));
                    break;
                }
                case SK_VELOCIMETER: {
                    V3<double> lin, ang, p;
                    Qt<double> q;
                    site_velocity(s, &lin, &ang);
                    parent_pose(s, &p, &q);
                    put(o, qrot_inv(qmul(q, lq), lin));
                    break;
                }
                case SK_MAGNETOMETER: {
                    V3<double> p; Qt<double> q;
                    parent_pose(s, &p, &q);
                    put(o, qrot_inv(qmul(q, lq), mk3(m.magnetic[0], m.magnetic[1], m.magnetic[2])));
                    break;
                }
                case SK_SUBTREE_COM: {
                    // subtree_com (sensor.rs:728-757): links in index order
                    // whose ancestor chain reaches the root link
                    V3<double> wsum = z3<double>();
                    double mass = 0.0;
                    for (int i = 0; i < m.n_links; ++i) {
                        int cur = i;
                        bool in = false;
                        while (cur >= 0) {
                            if (cur == s.link) { in = true; break; }
                            cur = m.links[cur].parent;
                        }
                        if (in) {
                            wsum = wsum + wd.pos(i) * m.links[i].mass;
                            mass = mass + m.links[i].mass;
                        }
                    }
                    put(o, mass == 0.0 ? z3<double>() : wsum / mass);
                    break;
                }
                case SK_FRAME_LINVEL: case SK_FRAME_ANGVEL: {
                    V3<double> lin, ang;
                    site_velocity(s, &lin, &ang);
                    put(o, s.kind == SK_FRAME_LINVEL ? lin : ang);
                    break;
                }
                default: break;
            }
        }
    }
};

}  // namespace newt
