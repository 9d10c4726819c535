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

#include "newt_ccd.hpp"
#include "newt_math.hpp"
#include "newt_types.hpp"

namespace newt {

// Rust's float `Sum`: a fold that starts at -0.0.
template <class R> struct Sum {
    R acc;
    NEWT_HD Sum() : acc(R(-0.0)) {}
    NEWT_HD void add(R x) { acc = acc + x; }
};

// face_along_direction (contact.rs:1209-1245): a vertex inside the box on
// every axis, assigned to the face along the dominant component of `dir`.
NEWT_HD inline bool face_along_direction(V3<double> lp, const double* h, V3<double> dir, double* pen,
                                         V3<double>* nl) {
    const double dx = h[0] - nabs(lp.x), dy = h[1] - nabs(lp.y), dz = h[2] - nabs(lp.z);
    if (dx < 0.0 || dy < 0.0 || dz < 0.0) return false;
    const double ax = nabs(dir.x), ay = nabs(dir.y), az = nabs(dir.z);
    double p;
    V3<double> n;
    if (ax >= ay && ax >= az) {
        const double sg = dir.x >= 0.0 ? 1.0 : -1.0;
        p = h[0] - sg * lp.x;
        n = mk3(sg, 0.0, 0.0);
    } else if (ay >= az) {
        const double sg = dir.y >= 0.0 ? 1.0 : -1.0;
        p = h[1] - sg * lp.y;
        n = mk3(0.0, sg, 0.0);
    } else {
        const double sg = dir.z >= 0.0 ? 1.0 : -1.0;
        p = h[2] - sg * lp.z;
        n = mk3(0.0, 0.0, sg);
    }
    if (p <= 0.0) return false;
    *pen = p;
    *nl = n;
    return true;
}

// Box-box vertex-vs-face manifold of the penalty pipeline
// (box_box_vertex_face, contact.rs:587-714): corners of A inside B (normal
// B's face along B->A, point moved onto B's surface), then corners of B
// inside A (A's face normal flipped, the corner itself); the four deepest,
// stable on ties, with the margin added.  Per-contact normals.
NEWT_HD inline int box_box_vertex_face(V3<double> pa, Qt<double> qa, const double* ha, V3<double> pb, Qt<double> qb,
                                       const double* hb, double margin, V3<double>* cp, double* pen,
                                       V3<double>* nrm) {
    const V3<double> dw = pa - pb;
    const V3<double> dib = qrot_inv(qb, dw);
    const V3<double> dia = qrot_inv(qa, -dw);
    const V3<double> dir_b = dot3(dib, dib) > 0.0 ? dib : mk3(0.0, 0.0, 1.0);
    const V3<double> dir_a = dot3(dia, dia) > 0.0 ? dia : mk3(-0.0, -0.0, -1.0);
    double cpen[16];
    V3<double> cpos[16], cn[16];
    int count = 0;
    for (int c = 0; c < 8; ++c) {  // corner signs in the order (-,-,-), (-,-,+), (-,+,-), ...
        const V3<double> la = mk3((c & 4) ? ha[0] : -ha[0], (c & 2) ? ha[1] : -ha[1], (c & 1) ? ha[2] : -ha[2]);
        const V3<double> wv = pa + qrot(qa, la);
        double p;
        V3<double> nl;
        if (face_along_direction(qrot_inv(qb, wv - pb), hb, dir_b, &p, &nl)) {
            const V3<double> nw = qrot(qb, nl);
            cpen[count] = p;
            cpos[count] = wv + nw * p;
            cn[count] = nw;
            ++count;
        }
    }
    for (int c = 0; c < 8; ++c) {
        const V3<double> lb = mk3((c & 4) ? hb[0] : -hb[0], (c & 2) ? hb[1] : -hb[1], (c & 1) ? hb[2] : -hb[2]);
        const V3<double> wv = pb + qrot(qb, lb);
        double p;
        V3<double> nl;
        if (face_along_direction(qrot_inv(qa, wv - pa), ha, dir_a, &p, &nl)) {
            cpen[count] = p;
            cpos[count] = wv;
            cn[count] = -qrot(qa, nl);
            ++count;
        }
    }
    int order[16];
    for (int i = 0; i < 16; ++i) order[i] = i;
    for (int i = 1; i < count; ++i)
        for (int j = i; j > 0 && cpen[order[j]] > cpen[order[j - 1]]; --j) {
            const int t = order[j - 1];
            order[j - 1] = order[j];
            order[j] = t;
        }
    const int take = count > 4 ? 4 : count;
    for (int i = 0; i < take; ++i) {
        cp[i] = cpos[order[i]];
        pen[i] = cpen[order[i]] + margin;
        nrm[i] = cn[order[i]];
    }
    return take;
}

template <class R> struct World {
    const ModelT& m;
    const Layout& L;
    Lane<R> ws;
    Lane<double> dw;
    Lane<int> iw;
    Lane<double> qpos, qvel, qacc, qfrc, fat;
    // Workspace offset of per-link external world wrenches (force, torque;
    // 6 per link) that aba applies, -1 for none; the sensors' post-step ABA
    // sets it.  Penalty contacts use their own wrenches (L.rbias).
    int ext_wr = -1;

    // ---------------- state access -----------------------------------------
    NEWT_HD R q(int k) const { return (R)qpos[k]; }
    NEWT_HD R qd(int k) const { return (R)qvel[k]; }

    // Kinematics and contact geometry are FP64 in both steps (dw): contact
    // depths are small differences of world coordinates.
    NEWT_HD V3<double> pos(int i) const { return {dw[L.poses + i * 7], dw[L.poses + i * 7 + 1], dw[L.poses + i * 7 + 2]}; }
    NEWT_HD Qt<double> ori(int i) const {
        return {dw[L.poses + i * 7 + 3], dw[L.poses + i * 7 + 4], dw[L.poses + i * 7 + 5], dw[L.poses + i * 7 + 6]};
    }
    NEWT_HD void set_pose(int i, V3<double> p, Qt<double> o) const {
        dw[L.poses + i * 7] = p.x; dw[L.poses + i * 7 + 1] = p.y; dw[L.poses + i * 7 + 2] = p.z;
        dw[L.poses + i * 7 + 3] = o.x; dw[L.poses + i * 7 + 4] = o.y; dw[L.poses + i * 7 + 5] = o.z;
        dw[L.poses + i * 7 + 6] = o.w;
    }
    NEWT_HD V3<double> kld3(int off) const { return {dw[off], dw[off + 1], dw[off + 2]}; }
    NEWT_HD void kst3(int off, V3<double> v) const { dw[off] = v.x; dw[off + 1] = v.y; dw[off + 2] = v.z; }
    NEWT_HD SM<double> kldm(int off) const { return {kld3(off), kld3(off + 3)}; }
    // Free bodies beside the trees: state at q_off / v_off of qpos / qvel.
    NEWT_HD V3<double> body_pos(int b) const {
        const int o = m.bodies[b].q_off;
        return mk3(qpos[o], qpos[o + 1], qpos[o + 2]);
    }
    NEWT_HD Qt<double> body_ori(int b) const {
        const int o = m.bodies[b].q_off;
        return Qt<double>{qpos[o + 3], qpos[o + 4], qpos[o + 5], qpos[o + 6]};
    }
    NEWT_HD V3<double> body_vel(int b) const {
        const int o = m.bodies[b].v_off;
        return mk3(qvel[o], qvel[o + 1], qvel[o + 2]);
    }
    NEWT_HD V3<double> body_omega(int b) const {  // body frame
        const int o = m.bodies[b].v_off;
        return mk3(qvel[o + 3], qvel[o + 4], qvel[o + 5]);
    }
    NEWT_HD V3<double> body_gravity(int b) const {
        const V3<double> g = mk3(m.gravity[0], m.gravity[1], m.gravity[2]);
        const double s = m.bodies[b].gravity_scale;
        return s == 1.0 ? g : g * s;
    }
    // A geom on a tree link (not a free body, not the static world).
    static NEWT_HD bool on_link(const GeomT& g) { return g.link >= 0 && g.body < 0; }
    NEWT_HD void kstm(int off, SM<double> v) const { kst3(off, v.w); kst3(off + 3, v.v); }
    NEWT_HD SF<double> kldf(int off) const { return {kld3(off), kld3(off + 3)}; }
    NEWT_HD void kstf(int off, SF<double> v) const { kst3(off, v.t); kst3(off + 3, v.f); }
    NEWT_HD Xf<double> kldx(int i) const {
        Xf<double> x;
        for (int k = 0; k < 9; ++k) x.E.m[k] = dw[L.xup + i * 12 + k];
        x.r = kld3(L.xup + i * 12 + 9);
        return x;
    }
    NEWT_HD void kstx(int i, const Xf<double>& x) const {
        for (int k = 0; k < 9; ++k) dw[L.xup + i * 12 + k] = x.E.m[k];
        kst3(L.xup + i * 12 + 9, x.r);
    }

    // ---------------- kinematics (tree.rs:771-858, world.rs:2777-2854) -----
    NEWT_HD void forward_kinematics() const {
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            V3<double> p;
            Qt<double> o;
            V3<double> pj = mk3((double)l.pj[0], (double)l.pj[1], (double)l.pj[2]);
            V3<double> cj = mk3((double)l.cj[0], (double)l.cj[1], (double)l.cj[2]);
            if (l.joint == JF_FREE) {
                int qo = l.q_off;
                p = mk3(qpos[qo], qpos[qo + 1], qpos[qo + 2]);
                o = {qpos[qo + 3], qpos[qo + 4], qpos[qo + 5], qpos[qo + 6]};
            } else if (l.joint == JF_FIXED) {
                V3<double> pp = z3<double>();
                Qt<double> po = qid<double>();
                if (l.parent >= 0) { pp = pos(l.parent); po = ori(l.parent); }
                Qt<double> pjq{(double)l.pjq[0], (double)l.pjq[1], (double)l.pjq[2], (double)l.pjq[3]};
                Qt<double> cjq{(double)l.cjq[0], (double)l.cjq[1], (double)l.cjq[2], (double)l.cjq[3]};
                V3<double> jp = pp + qrot(po, pj);
                Qt<double> jo = qmul(po, pjq);
                o = qmul(jo, qconj(cjq));
                p = jp - qrot(o, cj);
            } else {
                V3<double> pp = pos(l.parent);
                Qt<double> po = ori(l.parent);
                V3<double> ax = mk3((double)l.axis[0], (double)l.axis[1], (double)l.axis[2]);
                double qq = qpos[l.q_off];
                V3<double> jp = pp + qrot(po, pj);
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
    NEWT_HD void link_velocities() const {
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            V3<double> lin = z3<double>(), ang = z3<double>();
            if (l.parent < 0) {
                if (l.joint == JF_FREE) {
                    Qt<double> o = ori(i);
                    int vo = l.v_off;
                    ang = qrot(o, mk3(qvel[vo], qvel[vo + 1], qvel[vo + 2]));
                    lin = qrot(o, mk3(qvel[vo + 3], qvel[vo + 4], qvel[vo + 5]));
                }
            } else {
                V3<double> pv = kld3(L.lvel + l.parent * 6), pw = kld3(L.lvel + l.parent * 6 + 3);
                V3<double> cp = pos(i), pp = pos(l.parent);
                Qt<double> po = ori(l.parent);
                V3<double> pac = pv + cross3(pw, cp - pp);
                if (l.joint == JF_HINGE) {
                    V3<double> aw = qrot(po, mk3((double)l.axis[0], (double)l.axis[1], (double)l.axis[2]));
                    double r = qvel[l.v_off];
                    ang = pw + aw * r;
                    V3<double> jw = pp + qrot(po, mk3((double)l.pj[0], (double)l.pj[1], (double)l.pj[2]));
                    lin = pac + cross3(aw * r, cp - jw);
                } else if (l.joint == JF_SLIDE) {
                    V3<double> aw = qrot(po, mk3((double)l.axis[0], (double)l.axis[1], (double)l.axis[2]));
                    lin = pac + aw * qvel[l.v_off];
                    ang = pw;
                } else {  // fixed
                    lin = pac;
                    ang = pw;
                }
            }
            kst3(L.lvel + i * 6, lin);
            kst3(L.lvel + i * 6 + 3, ang);
        }
    }

    // geom_world_pose (geom.rs:887-898)
    NEWT_HD void geom_pose(const GeomT& g, V3<double>* p, Qt<double>* o) const {
        V3<double> gp = mk3((double)g.pos[0], (double)g.pos[1], (double)g.pos[2]);
        Qt<double> gq{(double)g.quat[0], (double)g.quat[1], (double)g.quat[2], (double)g.quat[3]};
        if (g.link < 0) { *p = gp; *o = gq; return; }
        Qt<double> lo = ori(g.link);
        *p = pos(g.link) + qrot(lo, gp);
        *o = qmul(lo, gq);
    }

    // geom_aabb (broadphase.rs:708-770) as (min, max).
    NEWT_HD void geom_aabb(const GeomT& g, V3<double> p, Qt<double> o, V3<double>* mn, V3<double>* mx) const {
        if (g.shape == SH_MESH) {  // transformed_vertices_aabb (broadphase.rs:780-791)
            const ccd::Mesh ms = mesh_of(g, p, o);
            *mn = ms.world(0);
            *mx = *mn;
            for (int k = 1; k < ms.nv; ++k) {
                const V3<double> w = ms.world(k);
                *mn = mk3(nmin(mn->x, w.x), nmin(mn->y, w.y), nmin(mn->z, w.z));
                *mx = mk3(nmax(mx->x, w.x), nmax(mx->y, w.y), nmax(mx->z, w.z));
            }
            return;
        }
        geom_aabb_of(g, p, o, mn, mx);
    }
    NEWT_HD ccd::Mesh mesh_of(const GeomT& g, V3<double> p, Qt<double> o) const {
        const MeshT& mt = m.meshes[g.mesh];
        return ccd::Mesh{m.mesh_v + 3 * mt.v_off, m.mesh_f + 3 * mt.f_off, mt.nv, mt.nf, p, o};
    }
    // Mesh pairs (dispatch_narrow_phase, contact.rs:2509-2590): mesh-mesh
    // through GJK/EPA (ccd::route_pair), mesh-plane and sphere-mesh in either
    // order; a swapped primitive's contacts keep the pair's geom order with
    // the normal negated.
    __host__ __device__ __noinline__ void mesh_pair(int* nc, int ia, const GeomT& ga, int ib, const GeomT& gb, double fr,
                                                    double margin, double gap) const {
        V3<double> pa, pb;
        Qt<double> qa, qb;
        geom_pose(ga, &pa, &qa);
        geom_pose(gb, &pb, &qb);
        ccd::Buf out;
        bool flip = false;
        if (ga.shape == SH_MESH && gb.shape == SH_MESH) {
            const ccd::Work wk{dw, iw, L.cv, L.cvi, {L.cf0, L.cf1}, {L.cfi0, L.cfi1}, L.ce, L.cpoly};
            ccd::mesh_mesh(mesh_of(ga, pa, qa), mesh_of(gb, pb, qb), margin, wk, out);
        } else if (ga.shape == SH_MESH && gb.shape == SH_PLANE) {
            ccd::mesh_plane(mesh_of(ga, pa, qa), pb, qb, margin, out);
        } else if (ga.shape == SH_PLANE) {
            ccd::mesh_plane(mesh_of(gb, pb, qb), pa, qa, margin, out);
            flip = true;
        } else if (gb.shape == SH_MESH) {
            ccd::sphere_mesh(pa, (double)ga.size[0], mesh_of(gb, pb, qb), margin, out);
        } else {
            ccd::sphere_mesh(pb, (double)gb.size[0], mesh_of(ga, pa, qa), margin, out);
            flip = true;
        }
        for (int k = 0; k < out.n; ++k)
            emit(nc, ia, ib, k, out.c[k].pos, flip ? -out.c[k].normal : out.c[k].normal, out.c[k].pen, fr, gap);
    }
    static NEWT_HD void geom_aabb_of(const GeomT& g, V3<double> p, Qt<double> o, V3<double>* mn, V3<double>* mx) {
        if (g.shape == SH_PLANE) {
            double inf = double(__builtin_inf());
            *mn = mk3(-inf, -inf, -inf);
            *mx = mk3(inf, inf, inf);
            return;
        }
        V3<double> e;
        if (g.shape == SH_SPHERE) {
            e = mk3((double)g.size[0], (double)g.size[0], (double)g.size[0]);
        } else if (g.shape == SH_BOX) {
            V3<double> x = abs3(qrot(o, mk3(double(1), double(0), double(0)))) * (double)g.size[0];
            V3<double> y = abs3(qrot(o, mk3(double(0), double(1), double(0)))) * (double)g.size[1];
            V3<double> z = abs3(qrot(o, mk3(double(0), double(0), double(1)))) * (double)g.size[2];
            e = (x + y) + z;
        } else {
            double r = (double)g.size[0];
            e = abs3(qrot(o, mk3(double(0), double(0), double(1)))) * (double)g.size[1] + mk3(r, r, r);
        }
        *mn = p - e;
        *mx = p + e;
    }

    // populate_broadphase_tree + swept_geom_bound + update_with_filter_and_tree
    // (world.rs:1254-1358, broadphase.rs:46-54, 320-346).  Needs current poses
    // and link velocities.
    NEWT_HD void populate(double dt, bool insert_all) const {
        for (int gi = 0; gi < m.n_geoms; ++gi) {
            const GeomT& g = m.geoms[gi];
            V3<double> sp;
            Qt<double> so;
            geom_pose(g, &sp, &so);
            V3<double> smn, smx, emn, emx;
            geom_aabb(g, sp, so, &smn, &smx);
            double ang_speed = double(0.0);
            V3<double> pivot = sp;
            if (g.body >= 0) {
                // a free body's end pose (world.rs:1317-1326): its own
                // linear velocity and body-frame angular velocity.
                const BodyT& b = m.bodies[g.body];
                V3<double> bp = mk3(qpos[b.q_off], qpos[b.q_off + 1], qpos[b.q_off + 2]);
                Qt<double> bo{qpos[b.q_off + 3], qpos[b.q_off + 4], qpos[b.q_off + 5], qpos[b.q_off + 6]};
                V3<double> bv = mk3(qvel[b.v_off], qvel[b.v_off + 1], qvel[b.v_off + 2]);
                V3<double> bw = mk3(qvel[b.v_off + 3], qvel[b.v_off + 4], qvel[b.v_off + 5]);
                V3<double> ep = bp + bv * dt;
                Qt<double> eo = qintegrate(bo, bw, dt);
                V3<double> gp = mk3((double)g.pos[0], (double)g.pos[1], (double)g.pos[2]);
                Qt<double> gq{(double)g.quat[0], (double)g.quat[1], (double)g.quat[2], (double)g.quat[3]};
                geom_aabb(g, ep + qrot(eo, gp), qmul(eo, gq), &emn, &emx);
                ang_speed = len3(qrot(bo, bw));
                pivot = bp;
            } else if (g.link >= 0) {
                V3<double> v = kld3(L.lvel + g.link * 6), w = kld3(L.lvel + g.link * 6 + 3);
                V3<double> lp = pos(g.link);
                Qt<double> lo = ori(g.link);
                V3<double> ep = lp + v * dt;
                Qt<double> eo = qintegrate(lo, qrot_inv(lo, w), dt);
                V3<double> gp = mk3((double)g.pos[0], (double)g.pos[1], (double)g.pos[2]);
                Qt<double> gq{(double)g.quat[0], (double)g.quat[1], (double)g.quat[2], (double)g.quat[3]};
                geom_aabb(g, ep + qrot(eo, gp), qmul(eo, gq), &emn, &emx);
                ang_speed = len3(w);
                pivot = lp;
            } else {
                emn = smn;
                emx = smx;
            }
            V3<double> bc = (smn + smx) * double(0.5);
            double br = len3((smx - smn) * double(0.5));
            double pr = len3(bc - pivot) + br;
            double sweep = (ang_speed > double(0.0) && nfinite(pr)) ? pr * ang_speed * dt : double(0.0);
            V3<double> mn = mk3(nmin(smn.x, emn.x), nmin(smn.y, emn.y), nmin(smn.z, emn.z));
            V3<double> mx = mk3(nmax(smx.x, emx.x), nmax(smx.y, emx.y), nmax(smx.z, emx.z));
            V3<double> s3 = mk3(sweep, sweep, sweep);
            mn = mn - s3;
            mx = mx + s3;
            double mg = (double)g.margin;
            V3<double> m3 = mk3(mg, mg, mg);
            mn = mn - m3;
            mx = mx + m3;
            int f0 = gi * 6;
            double fmnx = fat[f0], fmny = fat[f0 + 1], fmnz = fat[f0 + 2];
            double fmxx = fat[f0 + 3], fmxy = fat[f0 + 4], fmxz = fat[f0 + 5];
            bool inside = fmnx <= mn.x && fmny <= mn.y && fmnz <= mn.z && fmxx >= mx.x && fmxy >= mx.y && fmxz >= mx.z;
            if (!inside || insert_all) {
                V3<double> e = mx - mn;
                V3<double> fm = mk3(nmax(e.x * double(0.1), double(1.0e-4)), nmax(e.y * double(0.1), double(1.0e-4)), nmax(e.z * double(0.1), double(1.0e-4)));
                V3<double> a = mn - fm, b = mx + fm;
                fat[f0] = a.x; fat[f0 + 1] = a.y; fat[f0 + 2] = a.z;
                fat[f0 + 3] = b.x; fat[f0 + 4] = b.y; fat[f0 + 5] = b.z;
            }
        }
    }

    // ---------------- broadphase pairs + narrowphase -----------------------
    NEWT_HD int active_pairs() const {
        int n = 0;
        for (int k = 0; k < m.n_pairs; ++k) {
            int a = m.pairs[k].a * 6, b = m.pairs[k].b * 6;
            bool ov = fat[a] <= fat[b + 3] && fat[a + 3] >= fat[b] && fat[a + 1] <= fat[b + 4] &&
                      fat[a + 4] >= fat[b + 1] && fat[a + 2] <= fat[b + 5] && fat[a + 5] >= fat[b + 2];
            if (ov) iw[L.pairs + n++] = k;
        }
        return n;
    }

    NEWT_HD void emit(int* nc, int ga, int gb, int slot, V3<double> p, V3<double> n, double pen, double fr, double gap) const {
        int k = *nc;
        if (k >= m.cap_c) {  // beyond the compiled capacity: counted, not stored (a fault)
            *nc = k + 1;
            return;
        }
        iw[L.cids + 3 * k] = ga;
        iw[L.cids + 3 * k + 1] = gb;
        iw[L.cids + 3 * k + 2] = slot;
        int o = L.cgeo + 9 * k;
        kst3(o, p);
        kst3(o + 3, n);
        dw[o + 6] = pen;
        dw[o + 7] = fr;
        dw[o + 8] = gap;
        *nc = k + 1;
    }

    // One primitive with `prim` the non-plane geom (contact.rs:269-403);
    // `flip` relabels A/B and negates the normal (contact.rs:2579-2588).
    NEWT_HD void plane_primitive(int* nc, int ip, const GeomT& gp, int iq, const GeomT& gq, bool flip, double fr,
                                    double margin, double gap) const {
        V3<double> pp, qp;
        Qt<double> po, qo;
        geom_pose(gq, &pp, &po);  // plane pose
        geom_pose(gp, &qp, &qo);  // primitive pose
        V3<double> n = qrot(po, mk3(double(0), double(0), double(1)));
        int slot = 0;
        int ga = flip ? iq : ip, gb = flip ? ip : iq;
        V3<double> nn = flip ? -n : n;
        if (gp.shape == SH_SPHERE) {
            double radius = (double)gp.size[0];
            double raw = dot3(qp - pp, n) - radius;
            if (raw <= margin) emit(nc, ga, gb, slot++, qp - n * (radius + raw * double(0.5)), nn, margin - raw, fr, gap);
        } else if (gp.shape == SH_BOX) {
            double dist = dot3(qp - pp, n);
            for (int i = 0; i < 8; ++i) {
                V3<double> loc = mk3((i & 1) == 0 ? -(double)gp.size[0] : (double)gp.size[0], (i & 2) == 0 ? -(double)gp.size[1] : (double)gp.size[1],
                                (i & 4) == 0 ? -(double)gp.size[2] : (double)gp.size[2]);
                V3<double> co = qrot(qo, loc);
                double ld = dot3(n, co);
                double raw = dist + ld;
                if (raw > margin || ld > double(0.0)) continue;
                if (slot >= 4) continue;
                emit(nc, ga, gb, slot++, (qp + co) - n * (raw * double(0.5)), nn, margin - raw, fr, gap);
            }
        } else if (gp.shape == SH_CAPSULE) {
            double radius = (double)gp.size[0], hh = (double)gp.size[1];
            V3<double> aw = qrot(qo, mk3(double(0), double(0), double(1)));
            for (int e = 0; e < 2; ++e) {
                V3<double> end = e == 0 ? qp + aw * hh : qp - aw * hh;
                double raw = dot3(end - pp, n) - radius;
                if (raw <= margin) emit(nc, ga, gb, slot++, end - n * (radius + raw * double(0.5)), nn, margin - raw, fr, gap);
            }
        }
    }

    // legacy: the penalty pipeline's box-box (box_box, contact.rs:534-553),
    // vertex-vs-face first and the SAT manifold only when that is empty.
    template <bool legacy = false> NEWT_HD int narrowphase(int npairs) const {
        int nc = 0;
        for (int k = 0; k < npairs; ++k) {
            const PairT& pr = m.pairs[iw[L.pairs + k]];
            const GeomT& ga = m.geoms[pr.a];
            const GeomT& gb = m.geoms[pr.b];
            double fr = (double)ga.friction < (double)gb.friction ? (double)ga.friction : (double)gb.friction;
            double margin = (double)ga.margin > (double)gb.margin ? (double)ga.margin : (double)gb.margin;
            double gap = (double)ga.gap > (double)gb.gap ? (double)ga.gap : (double)gb.gap;
            // append_contacts (world.rs:2517-2528): the world's penetration
            // slop raises every contact's force-free gap to at least itself.
            if (m.slop > 0.0 && gap < m.slop) gap = m.slop;
            // (primitive, plane) in the given order, else swapped.
            if (ga.shape == SH_MESH || gb.shape == SH_MESH) {
                mesh_pair(&nc, pr.a, ga, pr.b, gb, fr, margin, gap);
            } else if (ga.shape == SH_BOX && gb.shape == SH_BOX) {
                V3<double> pa, pb, nrm, cp[4];
                Qt<double> qa, qb;
                double pen[4];
                geom_pose(ga, &pa, &qa);
                geom_pose(gb, &pb, &qb);
                if constexpr (legacy) {
                    V3<double> nv[4];
                    const int c = box_box_vertex_face(pa, qa, ga.size, pb, qb, gb.size, margin, cp, pen, nv);
                    for (int i = 0; i < c; ++i) emit(&nc, pr.a, pr.b, i, cp[i], nv[i], pen[i], fr, gap);
                    if (c > 0) continue;
                }
                const int c = box_box(pa, qa, ga.size, pb, qb, gb.size, margin, cp, pen, &nrm);
                for (int i = 0; i < c; ++i) emit(&nc, pr.a, pr.b, i, cp[i], nrm, pen[i], fr, gap);
            } else if (ga.shape != SH_PLANE && gb.shape != SH_PLANE) {
                V3<double> pa, pb, pt, nrm;
                Qt<double> qa, qb;
                double pen;
                geom_pose(ga, &pa, &qa);
                geom_pose(gb, &pb, &qb);
                if (round_pair(ga.shape == SH_CAPSULE, pa, qa, ga.size, gb.shape == SH_CAPSULE, pb, qb, gb.size,
                               margin, &pt, &nrm, &pen))
                    emit(&nc, pr.a, pr.b, 0, pt, nrm, pen, fr, gap);
            } else if (gb.shape == SH_PLANE) {
                plane_primitive(&nc, pr.a, ga, pr.b, gb, false, fr, margin, gap);
            } else {
                plane_primitive(&nc, pr.b, gb, pr.a, ga, true, fr, margin, gap);
            }
        }
        return nc;
    }

    // ---------------- Xup, CRB, Cholesky -----------------------------------
    // aba.rs transform (762-791); crb=true uses xup_for_link_hinge's -q
    // (tree.rs:1134-1140) instead of (0 - q): they differ only in the sign
    // of a zero.
    NEWT_HD Xf<double> xup(int i, bool crb) const {
        const LinkT& l = m.links[i];
        Xf<double> x;
        if (l.joint == JF_FIXED) {
#pragma unroll
            for (int k = 0; k < 9; ++k) x.E.m[k] = l.xfE[k];
            x.r = mk3(l.xfr[0], l.xfr[1], l.xfr[2]);
        } else if (l.joint == JF_HINGE) {
            V3<double> ax = unit3(mk3(l.axis[0], l.axis[1], l.axis[2]));
            double qq = qpos[l.q_off];
            double s, c;
            nsincos(crb ? (-qq) * double(0.5) : (double(0) - qq) * double(0.5), &s, &c);
            x.E = qmat(Qt<double>{ax.x * s, ax.y * s, ax.z * s, c});
            x.r = mk3(l.cj[0], l.cj[1], l.cj[2]) - mv(x.E, mk3(l.pj[0], l.pj[1], l.pj[2]));
        } else if (l.joint == JF_SLIDE) {
            for (int k = 0; k < 9; ++k) x.E.m[k] = (k % 4 == 0) ? double(1) : double(0);
            x.r = (mk3(l.cj[0], l.cj[1], l.cj[2]) - mk3(l.pj[0], l.pj[1], l.pj[2])) -
                  mk3(l.axis[0], l.axis[1], l.axis[2]) * qpos[l.q_off];
        } else {
            for (int k = 0; k < 9; ++k) x.E.m[k] = (k % 4 == 0) ? double(1) : double(0);
            x.r = z3<double>();
        }
        return x;
    }
    NEWT_HD SM<double> subspace(int i) const {
        const LinkT& l = m.links[i];
        V3<double> ax = mk3(l.axis[0], l.axis[1], l.axis[2]);
        if (l.joint == JF_HINGE) return {ax, cross3(mk3(l.cj[0], l.cj[1], l.cj[2]), ax)};
        return {z3<double>(), ax};
    }

    // dynamics::mass_matrix (dynamics.rs:474-578) into dw[M] (row-major nv*nv).
    NEWT_HD void mass_matrix() const {
        const int nv = m.tnv, n = m.n_links;
        for (int k = 0; k < nv * nv; ++k) dw[L.M + k] = 0.0;
        for (int i = 0; i < n; ++i) {
            for (int k = 0; k < 36; ++k) dw[L.ia + i * 36 + k] = m.links[i].si[k];
            kstx(i, xup(i, true));
        }
        // Several trees are one forest: every root (parent < 0) starts its
        // own tree, and M is block diagonal over the trees.
        for (int i = n - 1; i >= 1; --i) {
            int p = m.links[i].parent;
            if (p < 0) continue;
            double a[36], out[36];
            for (int k = 0; k < 36; ++k) a[k] = dw[L.ia + i * 36 + k];
            m6pullback(a, kldx(i), out);
            for (int k = 0; k < 36; ++k) dw[L.ia + p * 36 + k] = dw[L.ia + p * 36 + k] + out[k];
        }
        for (int i = 0; i < n; ++i) {
            if (m.links[i].parent >= 0 || m.links[i].joint != JF_FREE) continue;
            int o = m.links[i].v_off;
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) dw[L.M + (o + r) * nv + (o + c)] = dw[L.ia + i * 36 + r * 6 + c];
        }
        for (int i = 1; i < n; ++i) {
            const LinkT& l = m.links[i];
            if (l.joint == JF_FIXED || l.parent < 0) continue;
            int vi = l.v_off;
            SM<double> s = subspace(i);
            double a[36];
            for (int k = 0; k < 36; ++k) a[k] = dw[L.ia + i * 36 + k];
            SF<double> f = m6mul(a, s);
            double diag = sdot(s, f);
            diag = diag + l.armature;
            dw[L.M + vi * nv + vi] = (double)diag;
            int jc = i;
            while (m.links[jc].parent >= 0) {
                int j = m.links[jc].parent;
                f = xtforce(kldx(jc), f);
                jc = j;
                const LinkT& pl = m.links[j];
                int vj = pl.v_off;
                if (pl.joint == JF_HINGE || pl.joint == JF_SLIDE) {
                    double val = (double)sdot(subspace(j), f);
                    dw[L.M + vj * nv + vi] = val;
                    dw[L.M + vi * nv + vj] = val;
                } else if (pl.joint == JF_FREE) {
                    double fc[6] = {f.t.x, f.t.y, f.t.z, f.f.x, f.f.y, f.f.z};
                    for (int kk = 0; kk < 6; ++kk) {
                        dw[L.M + (vj + kk) * nv + vi] = (double)fc[kk];
                        dw[L.M + vi * nv + (vj + kk)] = (double)fc[kk];
                    }
                }
            }
        }
    }

    // Cholesky-Banachiewicz (dynamics.rs:602-622).  false when not PD.
    NEWT_HD bool cholesky(int src, int dst, bool add_damping, double dt) const {
        const int n = m.tnv;
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
    NEWT_HD void chol_solve(int fac, int vec, int res) const {
        const int n = m.tnv;
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
    // tau: dw[L.tau] generalized applied force; free_applied: whether the
    // free root reads it (disable_penalty_limits).  implicit: joint damping
    // folded into the pivots (Euler).  Result into dw[out].
    // Penalty contacts on the tree: the penalty solver with contact pairs.
    NEWT_HD bool penalty_contacts() const { return m.mode == 0 && m.n_pairs > 0; }

    // tree_wrenches_from_pairs (world.rs:2570-2624) with
    // apply_tree_contact_wrench (2633-2741) at the state whose poses and link
    // velocities are current: the legacy narrowphase over the step's pairs,
    // then per contact the spring-damper force on the tree's side(s) (the
    // static side only lends its zero velocity), world force and torque per
    // link at dw[L.rbias + 6 link].  Returns the contact count (above cap_c:
    // a capacity fault, nothing applied).
    NEWT_HD int tree_penalty_wrenches(int np) const {
        const int nc = narrowphase<true>(np);
        for (int k = 0; k < 6 * m.n_links; ++k) dw[L.rbias + k] = 0.0;
        if (nc > m.cap_c) return nc;
        for (int c = 0; c < nc; ++c) {
            const int o = L.cgeo + 9 * c;
            const GeomT& ga = m.geoms[iw[L.cids + 3 * c]];
            const GeomT& gb = m.geoms[iw[L.cids + 3 * c + 1]];
            const int la = ga.link, lb = gb.link;  // -1: static
            const double ma = la >= 0 ? m.links[la].mass * 1.0 : 0.0, mb = lb >= 0 ? m.links[lb].mass * 1.0 : 0.0;
            double m_eff;
            if (ma > 0.0 && mb > 0.0) m_eff = ma * mb / (ma + mb);
            else if (ma > 0.0) m_eff = ma;
            else if (mb > 0.0) m_eff = mb;
            else continue;
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
            const V3<double> cp = kld3(o), n = kld3(o + 3);
            V3<double> ra = z3<double>(), rb = z3<double>(), va = z3<double>(), vb = z3<double>();
            if (la >= 0) {  // point_velocity_generic (world.rs:2748-2775)
                ra = cp - pos(la);
                va = kld3(L.lvel + la * 6) + cross3(kld3(L.lvel + la * 6 + 3), ra);
            }
            if (lb >= 0) {
                rb = cp - pos(lb);
                vb = kld3(L.lvel + lb * 6) + cross3(kld3(L.lvel + lb * 6 + 3), rb);
            }
            const V3<double> vrel = va - vb;
            const double vn = dot3(vrel, n);
            const double pen = dw[o + 6], gap = dw[o + 8];
            const double pen_eff = m.slop > gap ? pen - m.slop : pen - gap;
            if (pen_eff <= 0.0) continue;
            const double f_raw = k * pen_eff - cd * vn;
            const double fn = f_raw > 0.0 ? f_raw : 0.0;
            if (fn <= 0.0) continue;
            const V3<double> ref = nabs(n.z) < 0.9 ? mk3(0.0, 0.0, 1.0) : mk3(1.0, 0.0, 0.0);
            const V3<double> t1 = unit3(cross3(ref, n));
            const V3<double> t2 = cross3(n, t1);
            const V3<double> vt = vrel - n * vn;
            const double vt1 = dot3(vt, t1), vt2 = dot3(vt, t2);
            const double cap = dw[o + 7] * fn;
            const double x1 = -cd * vt1, x2 = -cd * vt2;
            const double ft1 = x1 > cap ? cap : (x1 < -cap ? -cap : x1);
            const double ft2 = x2 > cap ? cap : (x2 < -cap ? -cap : x2);
            const V3<double> fa = (n * fn + t1 * ft1) + t2 * ft2;
            if (la >= 0) {
                kst3(L.rbias + 6 * la, kld3(L.rbias + 6 * la) + fa);
                kst3(L.rbias + 6 * la + 3, kld3(L.rbias + 6 * la + 3) + cross3(ra, fa));
            }
            if (lb >= 0) {
                const V3<double> fb = -fa;
                kst3(L.rbias + 6 * lb, kld3(L.rbias + 6 * lb) + fb);
                kst3(L.rbias + 6 * lb + 3, kld3(L.rbias + 6 * lb + 3) + cross3(rb, fb));
            }
        }
        return nc;
    }

    NEWT_HD void aba(bool free_applied, bool implicit, double dt, int out) const {
        const int n = m.n_links;
        for (int i = 0; i < n; ++i) {
            const LinkT& l = m.links[i];
            SM<double> v, c = smz<double>();
            if (l.joint == JF_FREE) {
                int o = l.v_off;
                kstx(i, xup(i, false));
                v = {mk3(qvel[o], qvel[o + 1], qvel[o + 2]), mk3(qvel[o + 3], qvel[o + 4], qvel[o + 5])};
            } else if (l.joint == JF_FIXED) {
                Xf<double> x = xup(i, false);
                kstx(i, x);
                v = xmotion(x, l.parent >= 0 ? kldm(L.v + l.parent * 6) : smz<double>());
            } else {
                Xf<double> x = xup(i, false);
                kstx(i, x);
                SM<double> s = subspace(i);
                kstm(L.S + i * 6, s);
                SM<double> sq = s * qvel[l.v_off];
                v = xmotion(x, kldm(L.v + l.parent * 6)) + sq;
                c = crm(v, sq);
            }
            kstm(L.v + i * 6, v);
            kstm(L.c + i * 6, c);
            // body-frame gravity + the external wrench (tree.rs:1049-1063):
            // the penalty contact wrench, else zero
            V3<double> gd = mk3(m.gravity[0], m.gravity[1], m.gravity[2]);
            V3<double> ef = z3<double>(), et = z3<double>();
            if (penalty_contacts()) {
                ef = kld3(L.rbias + 6 * i);
                et = kld3(L.rbias + 6 * i + 3);
            } else if (ext_wr >= 0) {
                ef = kld3(ext_wr + 6 * i);
                et = kld3(ext_wr + 6 * i + 3);
            }
            V3<double> fbd = qrot_inv(ori(i), ef + z3<double>() + gd * l.mass);
            V3<double> tbd = qrot_inv(ori(i), et + z3<double>());
            V3<double> fb = mk3((double)fbd.x, (double)fbd.y, (double)fbd.z);
            V3<double> tb = mk3((double)tbd.x, (double)tbd.y, (double)tbd.z);
            double a[36];
            for (int k = 0; k < 36; ++k) {
                a[k] = l.si[k];
                dw[L.ia + i * 36 + k] = a[k];
            }
            SF<double> pa = crf(v, m6mul(a, v)) - SF<double>{tb, fb};
            kstf(L.pa + i * 6, pa);
        }
        for (int i = n - 1; i >= 1; --i) {
            const LinkT& l = m.links[i];
            int p = l.parent;
            if (p < 0) continue;  // another tree's root
            double a[36], out6[36];
            for (int k = 0; k < 36; ++k) a[k] = dw[L.ia + i * 36 + k];
            Xf<double> x = kldx(i);
            SF<double> pa = kldf(L.pa + i * 6);
            SM<double> c = kldm(L.c + i * 6);
            SF<double> paf;
            if (l.joint == JF_FIXED) {
                paf = pa + m6mul(a, c);
            } else {
                SM<double> s = kldm(L.S + i * 6);
                SF<double> ias = m6mul(a, s);
                double dm = implicit ? dt * (l.damping + double(0.0)) : double(0.0);
                double d = (sdot(s, ias) + l.armature) + dm;
                SF<double> ps = pa + m6mul(a, c);
                double tauv = dw[L.tau + l.v_off];
                // joint_limit_scalar_force (tree.rs:1170-1207): the penalty
                // spring, while the constraint solver has not disabled it.
                double tl = 0.0;
                if (!free_applied && l.lim >= 0) {
                    const LimitT& lm = m.limits[l.lim];
                    const double q = qpos[l.q_off], qd = qvel[l.v_off];
                    if (q < lm.lo) {
                        tl = lm.stiffness * (lm.lo - q) + (qd < 0.0 ? -lm.damping * qd : 0.0);
                    } else if (q > lm.hi) {
                        tl = -lm.stiffness * (q - lm.hi) + (qd > 0.0 ? -lm.damping * qd : 0.0);
                    }
                }
                double fs = (((tauv + double(0.0)) - l.damping * qvel[l.v_off]) + tl) + double(-0.0);
                dw[L.fs + i] = fs;
                double u = fs - sdot(s, ps);
                double qdd = u / d;
                paf = ps + ias * qdd;
                double inv_d = double(1) / d;
                double iv[6] = {ias.t.x, ias.t.y, ias.t.z, ias.f.x, ias.f.y, ias.f.z};
                for (int r = 0; r < 6; ++r)
                    for (int cc = 0; cc < 6; ++cc) a[r * 6 + cc] = a[r * 6 + cc] - (iv[r] * iv[cc]) * inv_d;
                dw[L.d + i] = d;
            }
            m6pullback(a, x, out6);
            for (int k = 0; k < 36; ++k) dw[L.ia + p * 36 + k] = dw[L.ia + p * 36 + k] + out6[k];
            kstf(L.pa + p * 6, kldf(L.pa + p * 6) + xtforce(x, paf));
        }
        for (int k = 0; k < m.tnv; ++k) dw[out + k] = double(0);
        // Each root: a free root solves its 6x6 articulated system, a fixed
        // root does not accelerate.
        for (int ri = 0; ri < n; ++ri) {
        if (m.links[ri].parent >= 0) continue;
        SM<double> a0 = smz<double>();
        if (m.links[ri].joint == JF_FREE) {
            const LinkT& l = m.links[ri];
            int o = l.v_off;
            double fr[6];
            for (int k = 0; k < 6; ++k) fr[k] = double(0.0) + (free_applied ? dw[L.tau + o + k] : double(0.0));
            SF<double> pa0 = kldf(L.pa + ri * 6);
            SF<double> rhs = SF<double>{mk3(fr[0], fr[1], fr[2]), mk3(fr[3], fr[4], fr[5])} - pa0;
            double A[6][7];
            for (int r = 0; r < 6; ++r)
                for (int cc = 0; cc < 6; ++cc) A[r][cc] = dw[L.ia + ri * 36 + r * 6 + cc];
            double fd = l.free_damping;
            if (fd != double(0)) {
                SM<double> qr{mk3(qvel[o], qvel[o + 1], qvel[o + 2]), mk3(qvel[o + 3], qvel[o + 4], qvel[o + 5])};
                rhs = rhs - SF<double>{qr.w * fd, qr.v * fd};
                double dmr = implicit ? dt * fd : double(0.0);
                for (int k = 0; k < 6; ++k) A[k][k] = A[k][k] + dmr;
            }
            double rv[6] = {rhs.t.x, rhs.t.y, rhs.t.z, rhs.f.x, rhs.f.y, rhs.f.z};
            for (int r = 0; r < 6; ++r) A[r][6] = rv[r];
            // GMat6::solve (aba.rs:614-665): Gauss-Jordan, pivot = last row of
            // maximal |a| under the total order.
            for (int col = 0; col < 6; ++col) {
                int piv = col;
                for (int r = col + 1; r < 6; ++r)
                    if (!(nabs(A[piv][col]) > nabs(A[r][col]))) piv = r;
                if (piv != col)
                    for (int k = 0; k < 7; ++k) { double t = A[col][k]; A[col][k] = A[piv][k]; A[piv][k] = t; }
                double inv = double(1) / A[col][col];
                double prow[7];
                for (int k = 0; k < 7; ++k) prow[k] = A[col][k];
                for (int r = 0; r < 6; ++r) {
                    if (r == col) continue;
                    double factor = A[r][col] * inv;
                    for (int k = col; k < 7; ++k) A[r][k] = A[r][k] - factor * prow[k];
                }
            }
            double x6[6];
            for (int r = 0; r < 6; ++r) x6[r] = A[r][6] / A[r][r];
            a0 = {mk3(x6[0], x6[1], x6[2]), mk3(x6[3], x6[4], x6[5])};
            for (int k = 0; k < 6; ++k) dw[out + o + k] = x6[k];
        }
        kstm(L.a + ri * 6, a0);
        }
        for (int i = 1; i < n; ++i) {
            const LinkT& l = m.links[i];
            if (l.parent < 0) continue;
            Xf<double> x = kldx(i);
            SM<double> ap = xmotion(x, kldm(L.a + l.parent * 6));
            if (l.joint == JF_FIXED) {
                kstm(L.a + i * 6, ap);
                continue;
            }
            SM<double> prime = ap + kldm(L.c + i * 6);
            double a[36];
            for (int k = 0; k < 36; ++k) a[k] = dw[L.ia + i * 36 + k];
            SF<double> inner = m6mul(a, prime) + kldf(L.pa + i * 6);
            SM<double> s = kldm(L.S + i * 6);
            double qdd = (dw[L.fs + i] - sdot(s, inner)) / dw[L.d + i];
            kstm(L.a + i * 6, prime + s * qdd);
            dw[out + l.v_off] = qdd;
        }
    }

    // ---------------- contact rows (solver.rs:2805-3366) -------------------
    // Translational Jacobian of a world point on `link` (jacobian.rs:44-110),
    // into dw[L.jt + 3*slot] (FP64 geometry).
    NEWT_HD void point_jacobian(int link, V3<double> local) const {
        for (int s = 0; s < m.tnv; ++s) kst3(L.jt + 3 * s, z3<double>());
        V3<double> point = pos(link) + qrot(ori(link), local);
        const LinkT& tl = m.links[link];
        for (int k = 0; k < tl.chain_len; ++k) {
            int i = m.chain[tl.chain_off + k];
            const LinkT& l = m.links[i];
            if (l.joint == JF_FREE) {
                V3<double> r = point - pos(i);
                Qt<double> o = ori(i);
                V3<double> ex = qrot(o, mk3(1.0, 0.0, 0.0)), ey = qrot(o, mk3(0.0, 1.0, 0.0)), ez = qrot(o, mk3(0.0, 0.0, 1.0));
                kst3(L.jt + 3 * (l.v_off + 0), cross3(ex, r));
                kst3(L.jt + 3 * (l.v_off + 1), cross3(ey, r));
                kst3(L.jt + 3 * (l.v_off + 2), cross3(ez, r));
                kst3(L.jt + 3 * (l.v_off + 3), ex);
                kst3(L.jt + 3 * (l.v_off + 4), ey);
                kst3(L.jt + 3 * (l.v_off + 5), ez);
            } else if (l.joint == JF_HINGE) {
                Qt<double> po = ori(l.parent);
                V3<double> jw = pos(l.parent) + qrot(po, mk3(l.pj[0], l.pj[1], l.pj[2]));
                V3<double> aw = qrot(po, mk3(l.axis[0], l.axis[1], l.axis[2]));
                kst3(L.jt + 3 * l.v_off, cross3(aw, point - jw));
            } else if (l.joint == JF_SLIDE) {
                kst3(L.jt + 3 * l.v_off, qrot(ori(l.parent), mk3(l.axis[0], l.axis[1], l.axis[2])));
            }
        }
    }

    // tree_body_invweight0 (solver.rs:3719-3754), translational part, cached
    // per link for the step (it depends on the link only).  FP64.
    NEWT_HD double invweight(int link) const {
        if (iw[L.iwf + link]) return dw[L.iwc + link];
        point_jacobian(link, z3<double>());
        const int nv = m.tnv;
        double tr = 0.0;
        for (int axis = 0; axis < 3; ++axis) {
            for (int s = 0; s < nv; ++s) dw[L.dvec + s] = dw[L.jt + 3 * s + axis];
            chol_solve(L.Ld, L.dvec, L.dres);
            Sum<double> sm;
            for (int s = 0; s < nv; ++s) sm.add(dw[L.jt + 3 * s + axis] * dw[L.dres + s]);
            tr = tr + sm.acc;
        }
        double v = tr / 3.0;
        dw[L.iwc + link] = v;
        iw[L.iwf + link] = 1;
        return v;
    }

    // Builds rows, Delassus, PGS; leaves the generalized contact force in
    // ws[L.delta].  Returns false if a mass matrix is not positive definite.
    // Returns 0, or the fault: 3 a factor not positive definite, 4 a failed
    // Newton solve.
    NEWT_HD int solve_contacts(int nc, R dt, int* nrows_out) const {
        const int nv = m.tnv;
        const double dtd = m.dt;
        for (int s = 0; s < nv; ++s) dw[L.delta + s] = 0.0;
        *nrows_out = 0;
        // Active contacts (penetration beyond the force-free gap).
        bool any = false;
        for (int k = 0; k < nc; ++k)
            if (dw[L.cgeo + 9 * k + 6] - dw[L.cgeo + 9 * k + 8] > 0.0) any = true;
        if (!any) return 0;
        // Response factor (M + dt*B for Euler) and the diagApprox factor (M).
        mass_matrix();
        bool implicit = m.integrator != 2;
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            if (l.joint == JF_FREE) for (int k = 0; k < 6; ++k) dw[L.dvec + l.v_off + k] = l.free_damping;
            else if (l.joint == JF_HINGE || l.joint == JF_SLIDE) dw[L.dvec + l.v_off] = l.damping + 0.0;
        }
        if (!cholesky(L.M, L.Li, implicit && dtd != 0.0, dtd)) return 3;
        if (!cholesky(L.M, L.Ld, false, 0.0)) return 3;
        // Free bodies beside the trees: I_world^-1 = R I^-1 R^T per body
        // (solver.rs:2838-2845).  Their wrench accumulators (L.tbw) are
        // cleared by the step before the solve.
        for (int b = 0; b < m.n_bodies; ++b) {
            const M33<double> r = qmat(body_ori(b));
            const M33<double> iwm = m33_mul(m33_mul(r, m33_load(m.bodies[b].inv_inertia)), m33_tr(r));
            for (int k = 0; k < 9; ++k) dw[L.biw + 9 * b + k] = iwm.m[k];
        }
        // Free velocity: ABA at the current state, user forces only.
        for (int s = 0; s < nv; ++s) dw[L.tau + s] = qfrc[s];
        aba(true, implicit, dtd, L.qacc_free);
        for (int i = 0; i < m.n_links; ++i) iw[L.iwf + i] = 0;
        int nr = 0;
        bool frictionless = false;
        for (int k = 0; k < nc; ++k) {
            int o = L.cgeo + 9 * k;
            double pen = dw[o + 6] - dw[o + 8];
            if (pen <= 0.0) continue;
            const GeomT& ga = m.geoms[iw[L.cids + 3 * k]];
            const GeomT& gb = m.geoms[iw[L.cids + 3 * k + 1]];
            // contacts with no link side belong to the free-body solve
            if (!on_link(ga) && !on_link(gb)) continue;
            V3<double> cp = kld3(o), n = kld3(o + 3);
            double mu = ga.friction < gb.friction ? ga.friction : gb.friction;
            frictionless = frictionless || mu == 0.0;
            // tangent_basis (world.rs:3104-3111)
            V3<double> ref = nabs(n.z) < 0.9 ? mk3(0.0, 0.0, 1.0) : mk3(1.0, 0.0, 0.0);
            V3<double> t1 = unit3(cross3(ref, n));
            V3<double> t2 = cross3(n, t1);
            V3<double> dirs[4] = {n + t1 * mu, n - t1 * mu, n + t2 * mu, n - t2 * mu};
            for (int r = 0; r < 4; ++r) iw[L.rinfo + 3 * (nr + r)] = 0;
            for (int side = 0; side < 2; ++side) {
                const GeomT& g = side == 0 ? ga : gb;
                double sign = side == 0 ? 1.0 : -1.0;
                if (g.body >= 0) {
                    // world_jacobian_for_side, Body (solver.rs:3547-3565): the
                    // direction and its moment about the body's COM; the
                    // component is coded -(body + 1).
                    const V3<double> arm = cp - body_pos(g.body);
                    for (int r = 0; r < 4; ++r) {
                        int row = nr + r;
                        int cidx = iw[L.rinfo + 3 * row];
                        iw[L.rinfo + 3 * row + 1 + cidx] = -(g.body + 1);
                        iw[L.rinfo + 3 * row] = cidx + 1;
                        const int bo = L.bcomp + (row * 2 + cidx) * 6;
                        kst3(bo, dirs[r] * sign);
                        kst3(bo + 3, cross3(arm, dirs[r]) * sign);
                    }
                    continue;
                }
                if (g.link < 0) continue;
                V3<double> local = qrot_inv(ori(g.link), cp - pos(g.link));
                point_jacobian(g.link, local);
                for (int r = 0; r < 4; ++r) {
                    int row = nr + r;
                    int cidx = iw[L.rinfo + 3 * row];
                    iw[L.rinfo + 3 * row + 1 + cidx] = g.link;
                    iw[L.rinfo + 3 * row] = cidx + 1;
                    int co = L.coef + (row * 2 + cidx) * nv;
                    for (int s = 0; s < nv; ++s) ws[co + s] = (R)(dot3(kld3(L.jt + 3 * s), dirs[r]) * sign);
                }
            }
            // Per-row constants (solver.rs:3067-3120), FP64.
            double sr0 = ga.solref[0] <= gb.solref[0] ? ga.solref[0] : gb.solref[0];
            double sr1 = ga.solref[1] <= gb.solref[1] ? ga.solref[1] : gb.solref[1];
            if (ga.solref[0] < 0.0 || gb.solref[0] < 0.0) {
                if (!(ga.solref[0] < 0.0 && gb.solref[0] < 0.0)) {
                    const GeomT& gd = ga.solref[0] < 0.0 ? ga : gb;
                    sr0 = gd.solref[0];
                    sr1 = gd.solref[1];
                }
            }
            double dmin = ga.solimp[0] < gb.solimp[0] ? ga.solimp[0] : gb.solimp[0];
            double dmax = ga.solimp[1] < gb.solimp[1] ? ga.solimp[1] : gb.solimp[1];
            double width = ga.solimp[2] < gb.solimp[2] ? ga.solimp[2] : gb.solimp[2];
            double mid = ga.solimp[3];
            int power = ga.power;
            // effective_solimp (solver.rs:319-328)
            const double lo = 1.0e-4, hi = 0.9999;
            auto clampi = [&](double v) { if (v < lo) v = lo; if (v > hi) v = hi; return v; };
            dmin = clampi(dmin); dmax = clampi(dmax); mid = clampi(mid);
            width = width > 0.0 ? width : 0.0;
            if (power < 1) power = 1;
            auto impedance = [&](double position) {
                double x = (position - 0.0) / width;
                if (x < 0.0) x = -x;
                double y;
                if (x <= 0.0) y = 0.0;
                else if (x >= 1.0) y = 1.0;
                else if (x < mid) {
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
            double imp = impedance(-pen);
            // safe_solref (solver.rs:331-337) and reference_accel (353-384)
            double tc = sr0, dr = sr1;
            bool direct = tc < 0.0;
            if (!direct && tc < 2.0 * dtd) tc = 2.0 * dtd;
            double position = -pen;
            double dref = impedance(position);
            double bcoef, kcoef;
            if (direct) {
                bcoef = -dr / dmax;
                kcoef = -tc / (dmax * dmax);
            } else {
                bcoef = 2.0 / (dmax * tc);
                kcoef = dref / (dmax * dmax * tc * tc * dr * dr);
            }
            for (int r = 0; r < 4; ++r) {
                int row = nr + r;
                int ncomp = iw[L.rinfo + 3 * row];
                double tr = 0.0;
                Sum<double> vel, fvel;
                for (int cidx = 0; cidx < ncomp; ++cidx) {
                    const int who = iw[L.rinfo + 3 * row + 1 + cidx];
                    if (who < 0) {
                        // a body component (solver.rs:3624-3685, 3766-3771):
                        // 1/m into diagApprox, its current velocity and its
                        // gravity step along the row
                        const int b = -who - 1;
                        const int bo = L.bcomp + (row * 2 + cidx) * 6;
                        const V3<double> lin = kld3(bo), ang = kld3(bo + 3);
                        tr = tr + 1.0 / m.bodies[b].mass;
                        const double bv = dot3(lin, body_vel(b)) + dot3(ang, qrot(body_ori(b), body_omega(b)));
                        vel.add(bv + 0.0);
                        fvel.add(dot3(lin, body_gravity(b) * dtd) + 0.0);
                        continue;
                    }
                    int co = L.coef + (row * 2 + cidx) * nv;
                    tr = tr + invweight(who);
                    Sum<double> sv, sf;
                    for (int s = 0; s < nv; ++s) sv.add((double)ws[co + s] * qvel[s]);
                    for (int s = 0; s < nv; ++s) sf.add((double)ws[co + s] * (dw[L.qacc_free + s] * dtd));
                    vel.add(0.0 + sv.acc);
                    fvel.add(0.0 + sf.acc);
                }
                double diag_approx = tr * (1.0 + mu * mu);
                double base = (1.0 - imp) / imp * diag_approx;
                dw[L.rreg + row] = 2.0 * mu * mu * base;
                double refacc = -bcoef * vel.acc - kcoef * position;
                dw[L.rbias + row] = fvel.acc - refacc * dtd;
            }
            nr += 4;
        }
        *nrows_out = nr;
        if (nr == 0) return 0;
        // Delassus A = J (M+dtB)^-1 Jᵀ by columns (solver.rs:3045-3065): FP64
        // solves and dot products, stored in R.
        const int cap = m.cap_r;
        for (int col = 0; col < nr; ++col) {
            int ncomp = iw[L.rinfo + 3 * col];
            for (int s = 0; s < nv; ++s) dw[L.gres + s] = 0.0;
            // apply_world_impulse (solver.rs:3799-3836): a column's body
            // component moves its body by (d / m, R^T I_world^-1 (r x d)).
            int colb = -1;
            if (m.n_bodies > 0)
                for (int k = 0; k < 6; ++k) dw[L.bdel + k] = 0.0;
            for (int cidx = 0; cidx < ncomp; ++cidx) {
                const int who = iw[L.rinfo + 3 * col + 1 + cidx];
                if (who < 0) {
                    colb = -who - 1;
                    const int bo = L.bcomp + (col * 2 + cidx) * 6;
                    const V3<double> force = kld3(bo) * 1.0, torque = kld3(bo + 3) * 1.0;
                    M33<double> iwm;
                    for (int k = 0; k < 9; ++k) iwm.m[k] = dw[L.biw + 9 * colb + k];
                    const V3<double> av = mv(iwm, torque);
                    kst3(L.bdel, kld3(L.bdel) + force / m.bodies[colb].mass);
                    kst3(L.bdel + 3, kld3(L.bdel + 3) + qrot_inv(body_ori(colb), av));
                    continue;
                }
                int co = L.coef + (col * 2 + cidx) * nv;
                for (int s = 0; s < nv; ++s) dw[L.dvec + s] = (double)ws[co + s] * 1.0;
                chol_solve(L.Li, L.dvec, L.dres);
                for (int s = 0; s < nv; ++s) dw[L.gres + s] = dw[L.gres + s] + dw[L.dres + s];
            }
            for (int row = 0; row < nr; ++row) {
                int nrc = iw[L.rinfo + 3 * row];
                Sum<double> outer;
                for (int cidx = 0; cidx < nrc; ++cidx) {
                    const int who = iw[L.rinfo + 3 * row + 1 + cidx];
                    if (who < 0) {
                        // world_velocity_from_delta (solver.rs:3687-3714)
                        const int b = -who - 1;
                        const int bo = L.bcomp + (row * 2 + cidx) * 6;
                        const V3<double> dl = b == colb ? kld3(L.bdel) : z3<double>();
                        const V3<double> da = b == colb ? kld3(L.bdel + 3) : z3<double>();
                        const double bv = dot3(kld3(bo), dl) + dot3(kld3(bo + 3), qrot(body_ori(b), da));
                        outer.add(bv + 0.0);
                        continue;
                    }
                    int co = L.coef + (row * 2 + cidx) * nv;
                    Sum<double> sm;
                    for (int s = 0; s < nv; ++s) sm.add((double)ws[co + s] * dw[L.gres + s]);
                    outer.add(0.0 + sm.acc);
                }
                ws[L.A + row * cap + col] = (R)outer.acc;
            }
        }
        for (int row = 0; row < nr; ++row) {
            dw[L.rdiag + row] = (double)ws[L.A + row * cap + row] + dw[L.rreg + row];
            dw[L.f + row] = 0.0;
        }
        if (m.mode == 2) {
            if (!newton_contacts(nr, frictionless)) return 4;
        } else
        // Projected Gauss-Seidel (solver.rs:3269-3331, 3838-3861): Delassus in
        // R, impulses and residual accumulation in FP64.
        for (int it = 0; it < m.iterations; ++it)
            for (int i = 0; i < nr; ++i) {
                Sum<double> sm;
                for (int col = 0; col < nr; ++col) sm.add((double)ws[L.A + i * cap + col] * dw[L.f + col]);
                double fi = dw[L.f + i];
                double residual = (dw[L.rbias + i] + sm.acc) + dw[L.rreg + i] * fi;
                dw[L.f + i] = nmax(fi - residual / dw[L.rdiag + i], 0.0);
            }
        // Generalized force (solver.rs:3345-3364).
        for (int i = 0; i < nr; ++i) {
            double force = dw[L.f + i] / dtd;
            int ncomp = iw[L.rinfo + 3 * i];
            for (int cidx = 0; cidx < ncomp; ++cidx) {
                const int who = iw[L.rinfo + 3 * i + 1 + cidx];
                if (who < 0) {
                    // TreeContactSolution::body_wrenches (solver.rs:3345-3350)
                    const int bo = L.bcomp + (i * 2 + cidx) * 6, wo = L.tbw + 6 * (-who - 1);
                    kst3(wo, kld3(wo) + kld3(bo) * force);
                    kst3(wo + 3, kld3(wo + 3) + kld3(bo + 3) * force);
                    continue;
                }
                int co = L.coef + (i * 2 + cidx) * nv;
                for (int s = 0; s < nv; ++s) dw[L.delta + s] = dw[L.delta + s] + (double)ws[co + s] * force;
            }
        }
        return 0;
    }

    // ---------------- Newton on the tree contact rows ----------------------
    // The nonnegative Newton of newton.rs:275-556 over H = A + diag(reg),
    // lin = the row biases (solver.rs:3161-3241), every pyramidal row
    // projected onto f >= 0.  H[i][j] is A in R off the diagonal and
    // dw[rdiag] on it (A_ii + reg_i, plus the pivot below); the factor has
    // stride nr.  Vectors: x = dw[f], then g, d, Hv, a point, the left and
    // right points and slope of a line-search interval, a gradient.
    NEWT_HD double hval(int i, int j) const {
        return i == j ? dw[L.rdiag + i] : (double)ws[L.A + i * m.cap_r + j];
    }
    // cholesky (dynamics.rs:602-622) of H into dw[nfac]; false when a pivot
    // is not positive.
    NEWT_HD bool dense_cholesky(int n) const {
        const int fac = L.nfac;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                double sum = hval(i, j);
                for (int k = 0; k < j; ++k) sum -= dw[fac + i * n + k] * dw[fac + j * n + k];
                if (i == j) {
                    if (sum <= 0.0) return false;
                    dw[fac + i * n + i] = __builtin_sqrt(sum);
                } else {
                    dw[fac + i * n + j] = sum / dw[fac + j * n + j];
                }
            }
        return true;
    }
    // out = H v, each row summed in column order from 0.0 (newton.rs:603-612).
    NEWT_HD void hmul(int n, int v, int out) const {
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += hval(i, j) * dw[v + j];
            dw[out + i] = s;
        }
    }
    NEWT_HD double ndot(int n, int a, int b) const {  // Iterator::sum from -0.0
        Sum<double> s;
        for (int i = 0; i < n; ++i) s.add(dw[a + i] * dw[b + i]);
        return s.acc;
    }
    // dw[out] = max(x + alpha d, 0)
    NEWT_HD void naffine(int n, int x, int d, double alpha, int out) const {
        for (int i = 0; i < n; ++i) {
            double v = dw[x + i] + alpha * dw[d + i];
            dw[out + i] = v < 0.0 ? 0.0 : v;
        }
    }
    // quadratic_cost 0.5 pᵀHp + linᵀp, with Hp into hv
    NEWT_HD double ncost(int n, int p, int hv) const {
        hmul(n, p, hv);
        return 0.5 * ndot(n, p, hv) + ndot(n, L.rbias, p);
    }
    NEWT_HD bool newton_contacts(int n, bool frictionless) const {
        const int x = L.f, g = L.nvec, d = g + m.cap_r, hv = d + m.cap_r, pt = hv + m.cap_r;
        const int lp = pt + m.cap_r, rp = lp + m.cap_r, sl = rp + m.cap_r, gm = sl + m.cap_r;
        const int brk = L.nbrk;
        // A zero-friction contact repeats its normal row four times, so H is
        // only semidefinite: a solver-only diagonal pivot (solver.rs:3188-3203).
        if (!dense_cholesky(n)) {
            if (!frictionless) return false;
            double scale = 0.0;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) scale = nmax(scale, nabs(hval(i, j)));
            const double pivot = nmax(scale * 1.0e-6, 1.0e-7);
            for (int i = 0; i < n; ++i) dw[L.rdiag + i] = dw[L.rdiag + i] + pivot;
            if (!dense_cholesky(n)) return false;
        }
        const double tol = 1e-7;
        double cur = ncost(n, x, hv);
        for (int it = 0; it < m.iterations; ++it) {
            // g = H x + lin; d = -H⁻¹ g (cholesky_solve, dynamics.rs:626-648)
            hmul(n, x, g);
            for (int i = 0; i < n; ++i) dw[g + i] = dw[g + i] + dw[L.rbias + i];
            for (int i = 0; i < n; ++i) {
                double s = -dw[g + i];
                for (int k = 0; k < i; ++k) s -= dw[L.nfac + i * n + k] * dw[d + k];
                dw[d + i] = s / dw[L.nfac + i * n + i];
            }
            for (int i = n - 1; i >= 0; --i) {
                double s = dw[d + i];
                for (int k = i + 1; k < n; ++k) s -= dw[L.nfac + k * n + i] * dw[d + k];
                dw[d + i] = s / dw[L.nfac + i * n + i];
            }
            if (ndot(n, g, d) >= 0.0) break;
            // exact_pyramidal_line_search (newton.rs:400-556): breakpoints 0,
            // 1 and each -x_i / d_i inside (0, 1), sorted (stable) and
            // deduplicated within 1e-7; candidates are the breakpoints, then
            // per interval the minimiser of the quadratic along the
            // interval's slope; the first strictly cheapest wins.
            int nb = 2;
            dw[brk] = 0.0;
            dw[brk + 1] = 1.0;
            for (int i = 0; i < n; ++i) {
                const double di = dw[d + i];
                if (di == 0.0) continue;
                const double a = -dw[x + i] / di;
                if (a > 0.0 && a < 1.0) dw[brk + nb++] = a;
            }
            for (int i = 1; i < nb; ++i) {  // insertion sort: stable
                const double v = dw[brk + i];
                int j = i;
                while (j > 0 && v < dw[brk + j - 1]) {
                    dw[brk + j] = dw[brk + j - 1];
                    --j;
                }
                dw[brk + j] = v;
            }
            int nd = 0;
            for (int i = 0; i < nb; ++i) {
                const double b = dw[brk + i];
                if (nd == 0 || !(nabs(b - dw[brk + nd - 1]) <= 1e-7)) dw[brk + nd++] = b;
            }
            double best_a = 0.0, best_c = cur;
            for (int k = 0; k < nd; ++k) {
                const double a = dw[brk + k];
                naffine(n, x, d, a, pt);
                const double c = ncost(n, pt, hv);
                if (c < best_c) {
                    best_c = c;
                    best_a = a;
                }
            }
            for (int k = 0; k + 1 < nd; ++k) {
                const double lo = dw[brk + k], hi = dw[brk + k + 1];
                if (hi - lo <= 1e-7) continue;
                const double mid = 0.5 * (lo + hi);
                const double eps = (hi - lo) * 1e-4;
                const double la = nmax(mid - eps, lo), ra = nmin(mid + eps, hi);
                naffine(n, x, d, la, lp);
                naffine(n, x, d, ra, rp);
                const double span = ra - la;
                if (span <= 0.0) continue;
                const double inv = 1.0 / span;
                for (int i = 0; i < n; ++i) dw[sl + i] = (dw[rp + i] - dw[lp + i]) * inv;
                naffine(n, x, d, mid, pt);
                hmul(n, pt, gm);
                for (int i = 0; i < n; ++i) dw[gm + i] = dw[gm + i] + dw[L.rbias + i];
                hmul(n, sl, hv);
                const double den = ndot(n, sl, hv);
                const double num = ndot(n, gm, sl);
                const double a = den > 0.0 ? nmin(nmax(mid - num / den, lo), hi) : mid;
                naffine(n, x, d, a, pt);
                const double c = ncost(n, pt, hv);
                if (c < best_c) {
                    best_c = c;
                    best_a = a;
                }
            }
            if (best_a <= 0.0 || cur - best_c <= tol * (1.0 + nabs(cur))) break;
            for (int i = 0; i < n; ++i) {
                const double v = dw[x + i] + best_a * dw[d + i];
                dw[x + i] = v < 0.0 ? 0.0 : v;
            }
            cur = best_c;
        }
        return true;
    }

    // ---------------- one World::step --------------------------------------
    // Precondition: poses and link velocities describe the current state.
    // pre: broadphase pairs, contacts and the contact solve; returns a fault
    // code.
    NEWT_HD int pre() const {
        const R dt = (R)m.dt;
        const int nv = m.tnv;
        int np = active_pairs();
        int nc = 0;
        bool constraint = m.mode == 1 || m.mode == 2;
        for (int s = 0; s < nv; ++s) dw[L.delta + s] = 0.0;
        if (constraint) {
            nc = narrowphase(np);
            if (nc > m.cap_c) return 8;  // NEWT_ERR_CAPACITY
            int nr;
            const int st = solve_contacts(nc, dt, &nr);  // NEWT_ERR_NOT_PD, NEWT_ERR_NEWTON_FAIL
            if (st != 0) return st;
        }
        iw[L.iwf + m.n_links] = np;
        iw[L.iwf + m.n_links + 1] = nc;
        if (constraint)
            solve_limits();
        else
            for (int s = 0; s < nv; ++s) dw[L.lim + s] = 0.0;
        return 0;
    }

    // post: implicit-damping ABA with the contact force in dw[L.delta],
    // semi-implicit Euler, and the broadphase refresh on the new state.
    // solve_tree_limits (solver.rs:2509-2628) over the limit rows of
    // build_tree_solver_rows (2337-2373): one row per hinge/slide outside its
    // range, in link order; A from Cholesky solves of M; PGS with f >= 0.
    // The per-dof force into dw[L.lim] (zero without active rows).
    NEWT_HD void solve_limits() const {
        const int nv = m.tnv;
        const double dt = m.dt;
        for (int s = 0; s < nv; ++s) dw[L.lim + s] = 0.0;
        if (m.n_ranged == 0 || nv == 0 || !(dt > 0.0)) return;
        int n = 0;
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            if (l.lim < 0) continue;
            const LimitT& lt = m.limits[l.lim];
            const double q = qpos[l.q_off];
            double c, viol;
            if (q < lt.lo) {
                c = 1.0;
                viol = lt.lo - q;
            } else if (q > lt.hi) {
                c = -1.0;
                viol = q - lt.hi;
            } else {
                continue;
            }
            iw[L.lsl + n] = l.v_off;
            iw[L.llm + n] = l.lim;
            dw[L.lc + n] = c;
            dw[L.lvi + n] = viol;
            ++n;
        }
        if (n == 0) return;
        mass_matrix();
        if (!cholesky(L.M, L.Ld, false, 0.0)) return;  // no limit force (solver.rs:2529-2532)
        for (int i = 0; i < n; ++i) {
            for (int s = 0; s < nv; ++s) dw[L.dres + s] = 0.0;
            const int sl = iw[L.lsl + i];
            dw[L.dres + sl] = dw[L.dres + sl] + dw[L.lc + i];
            chol_solve(L.Ld, L.dres, L.ldv + i * nv);
        }
        // TreeRow::dot_response / dot_qdot (solver.rs:2214-2230)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                dw[L.lA + i * n + j] = 0.0 + dw[L.lc + i] * dw[L.ldv + j * nv + iw[L.lsl + i]];
        for (int i = 0; i < n; ++i) {
            const LimitT& lt = m.limits[iw[L.llm + i]];
            const SolImpE si = solimp_effective(lt.solimp, lt.power);
            const double viol = dw[L.lvi + i];
            const double vrow = 0.0 + dw[L.lc + i] * qvel[iw[L.lsl + i]];
            const double aref = reference_accel(-viol, vrow, lt.solref[0], lt.solref[1], si, dt);
            dw[L.lb + i] = -aref * dt;
            const double aii = dw[L.lA + i * n + i];
            const double d = impedance_at(si, -viol);
            const double r = d > 0.0 ? (1.0 - d) / d * aii : 0.0;
            dw[L.ldg + i] = aii + r;
            dw[L.lf + i] = 0.0;
        }
        for (int it = 0; it < m.iterations; ++it)
            for (int i = 0; i < n; ++i) {
                double ri = dw[L.lb + i];
                for (int j = 0; j < n; ++j) ri = ri + dw[L.lA + i * n + j] * dw[L.lf + j];
                const double extra = (dw[L.ldg + i] - dw[L.lA + i * n + i]) * dw[L.lf + i];
                ri = ri + extra;
                const double delta = -ri / dw[L.ldg + i];
                const double nf = dw[L.lf + i] + delta;
                dw[L.lf + i] = nf < 0.0 ? 0.0 : nf;
            }
        for (int i = 0; i < n; ++i) {
            const double fdt = dw[L.lf + i] / dt;
            const int sl = iw[L.lsl + i];
            dw[L.lim + sl] = dw[L.lim + sl] + dw[L.lc + i] * fdt;
        }
    }

    // tree_deriv (tree.rs:1485-1540) at the current state into dw[L.rdq]
    // (position rates); the velocity rates are the accelerations.
    NEWT_HD void deriv() const {
        for (int i = 0; i < m.n_links; ++i) {
            const LinkT& l = m.links[i];
            const int qo = l.q_off, vo = l.v_off;
            if (l.joint == JF_FREE) {
                Qt<double> o{qpos[qo + 3], qpos[qo + 4], qpos[qo + 5], qpos[qo + 6]};
                V3<double> vw = qrot(o, mk3(qvel[vo + 3], qvel[vo + 4], qvel[vo + 5]));
                Qt<double> dq = qderiv(o, mk3(qvel[vo], qvel[vo + 1], qvel[vo + 2]));
                dw[L.rdq + qo] = vw.x;
                dw[L.rdq + qo + 1] = vw.y;
                dw[L.rdq + qo + 2] = vw.z;
                dw[L.rdq + qo + 3] = dq.x;
                dw[L.rdq + qo + 4] = dq.y;
                dw[L.rdq + qo + 5] = dq.z;
                dw[L.rdq + qo + 6] = dq.w;
            } else if (l.joint == JF_HINGE || l.joint == JF_SLIDE) {
                dw[L.rdq + qo] = qvel[vo];
            }
        }
    }

    // The tree part of step_trees (world.rs:1826-1946) with
    // rk4_step_with_workspace (tree.rs:1234-1364): the contact force held
    // constant over four explicit ABA stages; the first stage's
    // acceleration is reported as qacc.
    // Penalty contacts are sampled again at every stage (world.rs:1913-1931);
    // returns a capacity fault code or 0.
    NEWT_HD int rk4(bool constraint) const {
        const int nv = m.tnv, nq = m.tnq;
        const double dt = m.dt;
        for (int k = 0; k < nq; ++k) dw[L.rq0 + k] = qpos[k];
        for (int s = 0; s < nv; ++s) dw[L.rv0 + s] = qvel[s];
        const double h[3] = {dt * 0.5, dt * 0.5, dt};
        for (int st = 0; st < 4; ++st) {
            forward_kinematics();
            if (penalty_contacts()) {
                link_velocities();
                const int nc = tree_penalty_wrenches(iw[L.iwf + m.n_links]);
                iw[L.iwf + m.n_links + 1] = nc;
                if (nc > m.cap_c) {
                    // Frozen at the start-of-step state, not a stage state.
                    for (int k = 0; k < nq; ++k) qpos[k] = dw[L.rq0 + k];
                    for (int s = 0; s < nv; ++s) qvel[s] = dw[L.rv0 + s];
                    return 8;  // NEWT_ERR_CAPACITY
                }
            }
            aba(constraint, false, dt, L.delta);
            if (st == 0)
                for (int s = 0; s < nv; ++s) qacc[s] = dw[L.delta + s];
            deriv();
            // ((k1 + 2 k2) + 2 k3) + k4, accumulated in stage order.
            for (int k = 0; k < nq; ++k) {
                const double d = dw[L.rdq + k];
                dw[L.raq + k] = st == 0 ? d : (st == 3 ? dw[L.raq + k] + d : dw[L.raq + k] + 2.0 * d);
            }
            for (int s = 0; s < nv; ++s) {
                const double d = dw[L.delta + s];
                dw[L.rav + s] = st == 0 ? d : (st == 3 ? dw[L.rav + s] + d : dw[L.rav + s] + 2.0 * d);
            }
            if (st < 3) {  // advance_tree_state (tree.rs:1551-1573)
                for (int k = 0; k < nq; ++k) qpos[k] = dw[L.rq0 + k] + dw[L.rdq + k] * h[st];
                for (int s = 0; s < nv; ++s) qvel[s] = dw[L.rv0 + s] + dw[L.delta + s] * h[st];
            }
        }
        const double w6 = dt * (1.0 / 6.0);
        for (int k = 0; k < nq; ++k) qpos[k] = dw[L.rq0 + k] + dw[L.raq + k] * w6;
        for (int s = 0; s < nv; ++s) qvel[s] = dw[L.rv0 + s] + dw[L.rav + s] * w6;
        for (int i = 0; i < m.n_links; ++i) {
            if (m.links[i].joint != JF_FREE) continue;
            const int qo = m.links[i].q_off;
            Qt<double> o = qnormalize(Qt<double>{qpos[qo + 3], qpos[qo + 4], qpos[qo + 5], qpos[qo + 6]});
            qpos[qo + 3] = o.x;
            qpos[qo + 4] = o.y;
            qpos[qo + 5] = o.z;
            qpos[qo + 6] = o.w;
        }
        return 0;
    }

    // With refresh false the caller refreshes the broadphase itself (after
    // stepping the free bodies beside the trees).
    NEWT_HD int post(bool refresh = true) const {
        const R dt = (R)m.dt;
        const int nv = m.tnv;
        bool constraint = m.mode == 1 || m.mode == 2;
        // step_trees_euler / step_trees (world.rs:1826-2051): the limit force
        // plus the contact force is added to qfrc_applied for the step and
        // taken back off afterwards.  dw[L.lim] keeps that delta.
        for (int s = 0; s < nv; ++s) {
            const double dl = dw[L.lim + s] + dw[L.delta + s];
            dw[L.lim + s] = dl;
            dw[L.tau + s] = qfrc[s] + dl;
        }
        if (m.integrator == 2) {
            const int f = rk4(constraint);
            if (f != 0) return f;
            for (int s = 0; s < nv; ++s) qfrc[s] = dw[L.tau + s] - dw[L.lim + s];
            if (refresh) {
                forward_kinematics();
                link_velocities();
                populate(dt, false);
            }
            return 0;
        }
        if (penalty_contacts()) {  // step_trees_euler: sampled once at the current state
            const int nc = tree_penalty_wrenches(iw[L.iwf + m.n_links]);
            iw[L.iwf + m.n_links + 1] = nc;
            if (nc > m.cap_c) return 8;  // NEWT_ERR_CAPACITY
        }
        aba(constraint, true, m.dt, L.delta);
        for (int s = 0; s < nv; ++s) qfrc[s] = dw[L.tau + s] - dw[L.lim + s];
        for (int s = 0; s < nv; ++s) qacc[s] = dw[L.delta + s];
        // euler_step_with_workspace (tree.rs:1407-1418) + integrate_tree_positions
        // (1422-1475).  Velocity and position accumulation is FP64 in both
        // steps.
        const double dtd = m.dt;
        for (int s = 0; s < nv; ++s) qvel[s] = qvel[s] + dw[L.delta + s] * dtd;
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
        if (refresh) {
            forward_kinematics();
            link_velocities();
            populate(dt, false);
        }
        return 0;
    }

    NEWT_HD int step() const {
        int st = pre();
        if (st != 0) return st;
        return post();
    }
};

}  // namespace newt
