// golden_scenes.cpp — the reference's golden scenes rebuilt on the oracle.
//
// TEST INFRASTRUCTURE ONLY.  Each function rebuilds one scene from the
// reference's golden tests in `float` and serializes the state exactly as
// the test does; tests/test_oracle_golden.py compares the bytes with the
// fixture files under tests/golden/.
#include "newt_world.hpp"

#include <cstring>
#include <string>

using namespace oracle;
using F = float;

namespace {

Link<F> mk_link(int parent, int joint, V3<F> pj, Q<F> pq, V3<F> cj, F mass, M3<F> inertia) {
    Link<F> l;
    l.parent = parent;
    l.joint = joint;
    l.pj_pos = pj;
    l.pj_quat = pq;
    l.cj_pos = cj;
    l.cj_quat = qidentity<F>();
    l.mass = mass;
    l.inertia = inertia;
    return l;
}
Link<F> hinge_link(int parent, V3<F> axis, V3<F> pj, V3<F> cj, F mass, M3<F> inertia) {
    Link<F> l = mk_link(parent, J_HINGE, pj, qidentity<F>(), cj, mass, inertia);
    l.axis = axis;
    return l;
}
void put(std::vector<uint8_t>& out, F v) {
    uint8_t b[4];
    std::memcpy(b, &v, 4);
    out.insert(out.end(), b, b + 4);
}
void put_u32(std::vector<uint8_t>& out, uint32_t v) {
    uint8_t b[4];
    std::memcpy(b, &v, 4);
    out.insert(out.end(), b, b + 4);
}
void snap_tree(std::vector<uint8_t>& out, const Tree<F>& t) {
    for (F v : t.q) put(out, v);
    for (F v : t.qdot) put(out, v);
}

// newt/tests/joints_golden.rs:23-135
std::vector<uint8_t> joints_chain_and_floating() {
    Tree<F> chain;
    chain.push_link(mk_link(-1, J_FIXED, vzero<F>(), qidentity<F>(), vzero<F>(), 1.0f, m3diag(1.0f, 1.0f, 1.0f)));
    const F masses[3] = {1.1f, 0.7f, 1.3f};
    const F lengths[3] = {0.6f, 0.4f, 0.5f};
    for (int i = 0; i < 3; ++i) {
        F l = lengths[i], m = masses[i];
        F ip = (1.0f / 12.0f) * m * l * l;
        V3<F> pa = i == 0 ? vzero<F>() : v3(0.0f, 0.0f, -lengths[i - 1] * 0.5f);
        chain.push_link(hinge_link(i, v3(1.0f, 0.0f, 0.0f), pa, v3(0.0f, 0.0f, l * 0.5f), m, m3diag(ip, ip, 1e-6f)));
    }
    chain.q[chain.q_offset[1]] = 0.4f;
    chain.q[chain.q_offset[2]] = -0.3f;
    chain.q[chain.q_offset[3]] = 0.2f;
    chain.qdot[chain.v_offset[2]] = 0.5f;

    Tree<F> fl;
    F hx = 0.3f, hy = 0.2f, hz = 0.15f, m0 = 2.0f;
    F hx2 = hx * hx, hy2 = hy * hy, hz2 = hz * hz;
    M3<F> box_i = m3diag((m0 / 3.0f) * (hy2 + hz2), (m0 / 3.0f) * (hx2 + hz2), (m0 / 3.0f) * (hx2 + hy2));
    Q<F> q0 = from_axis_angle(v3(1.0f, 0.4f, -0.3f), 0.6f);
    fl.push_link(mk_link(-1, J_FREE, v3(0.5f, -1.0f, 4.5f), q0, vzero<F>(), m0, box_i));
    F L = 0.8f, m1 = 0.7f;
    F ip = (1.0f / 12.0f) * m1 * L * L;
    V3<F> axis = normalize(v3(1.0f, 0.3f, 0.0f));
    fl.push_link(hinge_link(0, axis, v3(0.0f, 0.0f, -hz), v3(0.0f, 0.0f, L * 0.5f), m1, m3diag(ip, ip, 1e-4f)));
    fl.qdot[fl.v_offset[1]] = 4.0f;

    std::vector<uint8_t> out;
    snap_tree(out, chain);
    snap_tree(out, fl);
    V3<F> gc{0.0f, 0.0f, -9.81f}, gf{0.0f, 0.0f, 0.0f};
    AbaWs<F> w1, w2;
    F dt = 0.005f;
    for (int s = 0; s < 1000; ++s) {
        rk4_step(chain, gc, dt, w1);
        rk4_step(fl, gf, dt, w2);
        if (s == 99) { snap_tree(out, chain); snap_tree(out, fl); }
    }
    snap_tree(out, chain);
    snap_tree(out, fl);
    return out;
}

// newt/tests/joints_mixed_golden.rs
std::vector<uint8_t> joints_mixed() {
    Tree<F> t;
    t.push_link(mk_link(-1, J_FREE, v3(0.1f, -0.2f, 1.4f), from_axis_angle(normalize(v3(1.0f, 0.3f, -0.2f)), 0.4f),
                        vzero<F>(), 1.7f, m3diag(0.4f, 0.3f, 0.35f)));
    F w0[6] = {0.2f, 0.15f, -0.05f, 0.05f, -0.1f, 0.03f};
    for (int k = 0; k < 6; ++k) t.qdot[k] = w0[k];
    t.push_link(hinge_link(0, normalize(v3(0.7f, 0.2f, 0.15f)), v3(0.2f, 0.0f, -0.15f), v3(0.0f, 0.0f, 0.3f), 0.9f,
                           m3diag(0.08f, 0.08f, 0.005f)));
    t.q[t.q_offset[1]] = 0.4f;
    t.qdot[t.v_offset[1]] = -0.3f;
    Link<F> sl = mk_link(1, J_SLIDE, v3(0.0f, 0.0f, -0.3f), qidentity<F>(), vzero<F>(), 0.6f, m3diag(0.02f, 0.02f, 0.02f));
    sl.axis = normalize(v3(0.3f, 0.9f, -0.1f));
    t.push_link(sl);
    t.q[t.q_offset[2]] = 0.15f;
    t.qdot[t.v_offset[2]] = 0.25f;
    Link<F> bl = mk_link(2, J_BALL, v3(0.05f, 0.0f, -0.1f), qidentity<F>(), v3(0.0f, 0.0f, 0.25f), 0.5f, m3diag(0.03f, 0.03f, 0.004f));
    bl.damping = 0.05f;
    t.push_link(bl);
    Q<F> qb = renormalize(from_axis_angle(v3(1.0f, 0.0f, 0.0f), 0.35f) * from_axis_angle(v3(0.0f, 0.0f, 1.0f), 0.2f));
    int o = t.q_offset[3];
    t.q[o] = qb.x; t.q[o + 1] = qb.y; t.q[o + 2] = qb.z; t.q[o + 3] = qb.w;
    int vo = t.v_offset[3];
    t.qdot[vo] = 0.3f; t.qdot[vo + 1] = 0.6f; t.qdot[vo + 2] = -0.4f;

    std::vector<uint8_t> out;
    auto ser = [&]() {
        put_u32(out, (uint32_t)t.q.size());
        for (F v : t.q) put(out, v);
        put_u32(out, (uint32_t)t.qdot.size());
        for (F v : t.qdot) put(out, v);
    };
    ser();
    AbaWs<F> ws;
    V3<F> g{0.0f, 0.0f, -9.81f};
    const int deltas[3] = {100, 400, 500};
    for (int d : deltas) {
        for (int s = 0; s < d; ++s) rk4_step(t, g, 0.005f, ws);
        ser();
    }
    return out;
}

M3<F> box_inertia(F mass, F hx, F hy, F hz) {
    F hx2 = hx * hx, hy2 = hy * hy, hz2 = hz * hz;
    return m3diag((mass / 3.0f) * (hy2 + hz2), (mass / 3.0f) * (hx2 + hz2), (mass / 3.0f) * (hx2 + hy2));
}

// newt/tests/dynamics_golden.rs
std::vector<uint8_t> dynamics_mass_bias() {
    Tree<F> t;
    t.push_link(mk_link(-1, J_FIXED, vzero<F>(), qidentity<F>(), vzero<F>(), 1.0f, m3diag(1.0f, 1.0f, 1.0f)));
    Link<F> h = hinge_link(0, normalize(v3(1.0f, 0.3f, -0.2f)), v3(0.05f, 0.0f, -0.1f), v3(-0.15f, 0.0f, 0.3f), 0.9f,
                           box_inertia(0.9f, 0.15f, 0.1f, 0.3f));
    h.armature = 0.02f;
    t.push_link(h);
    Link<F> s = mk_link(1, J_SLIDE, v3(0.0f, 0.0f, -0.3f), qidentity<F>(), v3(0.1f, -0.2f, 0.15f), 0.7f,
                        box_inertia(0.7f, 0.1f, 0.2f, 0.15f));
    s.axis = normalize(v3(0.2f, 1.0f, 0.4f));
    s.armature = 0.05f;
    t.push_link(s);
    Link<F> b = mk_link(2, J_BALL, v3(0.1f, -0.2f, -0.15f), qidentity<F>(), v3(-0.05f, 0.1f, 0.2f), 0.5f,
                        m3diag(0.02f, 0.03f, 0.04f));
    b.armature = 0.01f;
    t.push_link(b);
    t.q[t.q_offset[1]] = 0.35f;
    t.qdot[t.v_offset[1]] = 0.6f;
    t.q[t.q_offset[2]] = 0.12f;
    t.qdot[t.v_offset[2]] = -0.4f;
    Q<F> qb = renormalize(from_axis_angle(normalize(v3(0.4f, 1.0f, -0.3f)), 0.55f));
    int o = t.q_offset[3];
    t.q[o] = qb.x; t.q[o + 1] = qb.y; t.q[o + 2] = qb.z; t.q[o + 3] = qb.w;
    int vo = t.v_offset[3];
    t.qdot[vo] = 0.3f; t.qdot[vo + 1] = -0.5f; t.qdot[vo + 2] = 0.7f;
    std::vector<F> m = mass_matrix(t);
    std::vector<F> qdd(t.nv(), 0.0f);
    std::vector<std::pair<V3<F>, V3<F>>> ext(t.links.size(), {vzero<F>(), vzero<F>()});
    std::vector<F> bias = inverse_dynamics(t, qdd, v3(0.0f, 0.0f, -9.81f), ext);
    std::vector<uint8_t> out;
    for (F v : m) put(out, v);
    for (F v : bias) put(out, v);
    return out;
}

Geom<F> static_plane(F friction) {  // Geom::static_plane(ZERO, Z, friction), geom.rs:465
    Geom<F> g;
    g.shape = S_PLANE;
    g.pos = vzero<F>();
    g.quat = qidentity<F>();  // quat_align_z_to(+Z) = identity
    g.friction = friction;
    return g;
}
Geom<F> box_on_link(int tree, int link, V3<F> he, V3<F> off, Q<F> q, F friction) {  // geom.rs:604
    Geom<F> g;
    g.shape = S_BOX;
    g.tree = tree;
    g.link = link;
    g.size = he;
    g.pos = off;
    g.quat = q;
    g.friction = friction;
    return g;
}

// newt/tests/solver_golden.rs:93-144 + record_tree (272-297), PGS, RK4.
std::vector<uint8_t> solver_tree(int mode = M_PGS) {
    World<F> w;
    w.dt = 0.005f;
    w.gravity = v3(0.0f, 0.0f, -9.81f);
    w.mode = mode;
    w.iterations = 30;
    w.cone = C_PYRAMIDAL;
    w.integrator = I_RK4;
    Geom<F> plane = static_plane(0.7f);
    plane.solref = {0.05f, 1.5f};
    w.add_geom(plane);
    Tree<F> t;
    Q<F> ro = from_axis_angle(normalize(v3(1.0f, 0.4f, -0.2f)), 0.35f);
    t.push_link(mk_link(-1, J_FREE, v3(0.13f, -0.07f, 0.78f), ro, vzero<F>(), 1.2f, m3diag(0.08f, 0.09f, 0.1f)));
    t.push_link(hinge_link(0, v3(0.0f, 1.0f, 0.0f), v3(0.0f, 0.0f, -0.3f), v3(0.0f, 0.0f, 0.22f), 0.75f,
                           m3diag(0.02f, 0.025f, 0.018f)));
    t.q[t.q_offset[1]] = 0.37f;
    t.qdot[t.v_offset[1]] = -0.25f;
    int ti = w.add_tree(t, true);
    w.add_geom(box_on_link(ti, 0, v3(0.22f, 0.17f, 0.19f), v3(0.08f, -0.06f, 0.0f), from_axis_angle(v3(0.0f, 1.0f, 0.0f), 0.2f), 0.7f));
    w.add_geom(box_on_link(ti, 1, v3(0.14f, 0.12f, 0.24f), v3(0.0f, 0.0f, -0.18f), from_axis_angle(v3(1.0f, 0.0f, 0.0f), -0.15f), 0.7f));
    std::vector<uint8_t> out;
    auto snap = [&]() { for (auto& tr : w.trees) snap_tree(out, tr); };
    snap();
    for (int s = 0; s < 1000; ++s) {
        w.step();
        if (s == 49 || s == 199) snap();
    }
    snap();
    return out;
}

// Body state as the reference's golden tests serialize it: position,
// orientation (x, y, z, w), linear velocity, body angular velocity.
void snap_bodies(std::vector<uint8_t>& out, const World<F>& w) {
    for (const Body<F>& b : w.bodies) {
        const F v[13] = {b.p.x, b.p.y, b.p.z, b.q.x, b.q.y, b.q.z, b.q.w, b.v.x, b.v.y, b.v.z, b.w.x, b.w.y, b.w.z};
        for (F x : v) put(out, x);
    }
}
Body<F> solid_box(F mass, V3<F> h, V3<F> p, Q<F> q) { return make_body(mass, solid_box_inertia(mass, h), p, q); }
Geom<F> box_on_body(int body, V3<F> he, F friction) {  // Geom::r#box (geom.rs:575-600)
    Geom<F> g;
    g.shape = S_BOX;
    g.body = body;
    g.size = he;
    g.friction = friction;
    return g;
}
std::vector<uint8_t> record_at(World<F>& w, std::initializer_list<int> marks) {
    std::vector<uint8_t> out;
    snap_bodies(out, w);
    int done = 0;
    for (int m : marks) {
        for (; done < m; ++done) w.step();
        snap_bodies(out, w);
    }
    return out;
}

// newt/tests/golden.rs:18-93: three tumbling boxes, RK4, no contacts.
std::vector<uint8_t> tumbling_3_body() {
    World<F> w;
    w.dt = 0.005f;
    w.gravity = v3(0.0f, 0.0f, -9.81f);
    Body<F> b0 = solid_box(1.0f, v3(0.4f, 0.3f, 0.2f), v3(0.0f, 0.0f, 5.0f), qidentity<F>());
    b0.v = v3(1.0f, 0.5f, 3.0f);
    b0.w = v3(1.0f, 2.0f, 0.5f);
    Body<F> b1 = solid_box(2.0f, v3(0.3f, 0.3f, 0.3f), v3(1.5f, -1.0f, 4.0f), from_axis_angle(v3(1.0f, 0.0f, 0.0f), 0.3f));
    b1.v = v3(-0.5f, 1.5f, 2.0f);
    b1.w = v3(0.2f, 3.5f, 0.1f);
    Body<F> b2 = solid_box(0.5f, v3(0.5f, 0.2f, 0.1f), v3(-1.0f, 2.0f, 6.0f), from_axis_angle(v3(0.0f, 1.0f, 1.0f), 0.7f));
    b2.v = v3(2.0f, -1.0f, 0.5f);
    b2.w = v3(0.05f, 0.1f, 4.0f);
    w.add_body(b0);
    w.add_body(b1);
    w.add_body(b2);
    return record_at(w, {100, 1000});
}

// newt/tests/integrators_golden.rs:10-63: one body per integrator.
std::vector<uint8_t> integrator_golden(int integrator) {
    World<F> w;
    w.integrator = integrator;
    w.dt = 0.005f;
    w.gravity = v3(0.0f, 0.0f, -9.81f);
    Body<F> b = make_body(1.25f, m3diag(0.7f, 1.3f, 2.1f), v3(0.4f, -0.7f, 2.0f),
                          from_axis_angle(v3(1.0f, 0.2f, -0.4f), 0.6f));
    b.v = v3(0.8f, -0.4f, 1.2f);
    b.w = v3(0.4f, 1.1f, -0.3f);
    w.add_body(b);
    return record_at(w, {20, 100});
}

// newt/tests/solver_golden.rs:27-79 + record (170-190): PGS 30, pyramidal, RK4.
std::vector<uint8_t> solver_stack(int mode = M_PGS) {
    World<F> w;
    w.dt = 0.005f;
    w.gravity = v3(0.0f, 0.0f, -9.81f);
    w.mode = mode;
    w.iterations = 30;
    w.cone = C_PYRAMIDAL;
    SolRef<F> sr{0.05f, 1.5f};
    Geom<F> plane = static_plane(0.8f);
    plane.solref = sr;
    w.add_geom(plane);
    V3<F> half{0.25f, 0.25f, 0.25f};
    for (int k = 0; k < 3; ++k) {
        V3<F> off = v3(0.01f * (F)k, 0.0f, 0.25f + 0.5f * (F)k);
        int idx = w.add_body(solid_box(1.0f, half, off, qidentity<F>()));
        Geom<F> g = box_on_body(idx, half, 0.8f);
        g.solref = sr;
        w.add_geom(g);
    }
    return record_at(w, {50, 200, 1000});
}
std::vector<uint8_t> solver_incline(int mode = M_PGS) {
    World<F> w;
    w.dt = 0.005f;
    F angle = 32.0f * 3.14159265358979323846f / 180.0f;
    F g = 9.81f;
    w.gravity = v3(g * nsin(angle), 0.0f, -g * ncos(angle));
    w.mode = mode;
    w.iterations = 30;
    w.cone = C_PYRAMIDAL;
    SolRef<F> sr{0.05f, 1.5f};
    Geom<F> plane = static_plane(0.5f);
    plane.solref = sr;
    w.add_geom(plane);
    V3<F> half{0.2f, 0.21f, 0.19f};
    int bi = w.add_body(solid_box(1.0f, half, v3(0.01f, 0.0f, 0.2f), qidentity<F>()));
    Geom<F> bg = box_on_body(bi, half, 0.5f);
    bg.solref = sr;
    w.add_geom(bg);
    return record_at(w, {50, 200, 1000});
}

}  // namespace

// newt/tests/sensors_battery.rs:1056-1152: a pendulum tree and a tumbling
// free body in one world (World::new: RK4, penalty solver, no geoms), one
// sensor of each of six kinds; sensordata after steps 1, 100 and 500.
std::vector<uint8_t> sensors_battery() {
    World<F> w;
    w.dt = 0.005f;
    w.gravity = v3(0.0f, 0.0f, -9.81f);
    Tree<F> tree;  // pendulum_tree (sensors_battery.rs:376-395)
    tree.push_link(mk_link(-1, J_FIXED, vzero<F>(), qidentity<F>(), vzero<F>(), 1.0f, m3diag(1.0f, 1.0f, 1.0f)));
    tree.push_link(hinge_link(0, v3(1.0f, 0.0f, 0.0f), vzero<F>(), v3(0.0f, 0.0f, 1.0f), 1.0f,
                              m3diag(1e-3f, 1e-3f, 1e-3f)));
    tree.q[tree.q_offset[1]] = 0.4f;      // set_hinge_angle(1, 0.4)
    tree.qdot[tree.v_offset[1]] = -0.2f;  // set_hinge_rate(1, -0.2)
    w.add_tree(tree, true);
    Body<F> b = make_body(1.0f, m3diag(1.0f, 2.0f, 3.0f), v3(2.0f, 0.0f, 3.0f),
                          from_axis_angle(v3(1.0f, 0.4f, -0.3f), 0.7f));
    b.w = v3(0.6f, 0.9f, -0.4f);
    b.v = v3(-0.1f, 0.2f, 0.05f);
    w.add_body(b);
    using W = World<F>;
    w.add_sensor(W::SK_JPOS, 0, 1, -1, -1, vzero<F>(), qidentity<F>());
    w.add_sensor(W::SK_JVEL, 0, 1, -1, -1, vzero<F>(), qidentity<F>());
    w.add_sensor(W::SK_FPOS, 0, 1, -1, -1, v3(0.0f, 0.0f, -0.5f), qidentity<F>());
    w.add_sensor(W::SK_FQUAT, 0, 1, -1, -1, vzero<F>(), qidentity<F>());
    w.add_sensor(W::SK_GYRO, 0, -1, 0, -1, v3(0.1f, 0.2f, 0.3f), qidentity<F>());
    w.add_sensor(W::SK_ACCEL, 0, -1, 0, -1, v3(0.1f, 0.2f, 0.3f), qidentity<F>());
    std::vector<uint8_t> out;
    int done = 0;
    for (int at : {1, 100, 500}) {
        for (; done < at; ++done) w.step();
        for (F v : w.sensordata) put(out, v);
    }
    return out;
}

extern "C" int oracle_golden_bytes(const char* name, uint8_t* out, int cap) {
    std::vector<uint8_t> b;
    std::string n(name);
    if (n == "joints_chain_and_floating") b = joints_chain_and_floating();
    else if (n == "joints_mixed") b = joints_mixed();
    else if (n == "dynamics_mass_bias") b = dynamics_mass_bias();
    else if (n == "solver_tree") b = solver_tree();
    else if (n == "solver_newton_tree") b = solver_tree(M_NEWTON);
    else if (n == "tumbling_3_body") b = tumbling_3_body();
    else if (n == "integrator_euler") b = integrator_golden(I_EULER);
    else if (n == "integrator_rk4") b = integrator_golden(I_RK4);
    else if (n == "integrator_implicitfast") b = integrator_golden(I_IMPLICITFAST);
    else if (n == "solver_stack") b = solver_stack();
    else if (n == "solver_incline") b = solver_incline();
    else if (n == "solver_newton_stack") b = solver_stack(M_NEWTON);
    else if (n == "solver_newton_incline") b = solver_incline(M_NEWTON);
    else if (n == "sensors_battery") b = sensors_battery();
    else return -1;
    if ((int)b.size() > cap) return -2;
    std::memcpy(out, b.data(), b.size());
    return (int)b.size();
}
