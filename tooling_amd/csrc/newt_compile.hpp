// newt_compile.hpp — host-side compilation of a model description
// (include/newt_gpu.h records) into the flat tables of newt_types.hpp, and
// the per-world workspace layout.  No device calls: newt_gpu.hip uploads the
// tables; the same code serves any host-side consumer of the tables.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/newt_gpu.h"
#include "newt_math.hpp"
#include "newt_types.hpp"

namespace newt {

struct HostModel {
    newt_options opt{};
    std::vector<std::vector<newt_link_desc>> trees;
    std::vector<int> tree_collide;
    std::vector<newt_geom_desc> geoms;
    ModelT dev{};  // pointers into the vectors below
    std::vector<LinkT> links;
    std::vector<GeomT> hgeoms;
    std::vector<PairT> pairs;
    std::vector<int> chain;
    std::vector<int> level_off, level_links, level_nrank, dof_link;
    std::vector<LimitT> limits;
    std::vector<double> q_default;
    std::vector<double> v_default;  // zero for trees; the bodies' initial velocities
    std::vector<newt_body_desc> bodies;
    std::vector<BodyT> hbodies;
    int contact_cap = 0;  // most contacts per world (0: the sum of the pairs' maxima)
    std::vector<newt_sensor_desc> sensors;
    std::vector<SensorT> hsensors;
    double magnetic[3] = {0.0, 0.0, 0.0};
    std::vector<MeshT> meshes;  // World::meshes, as added
    std::vector<double> mesh_v;
    std::vector<int> mesh_f;
};



inline newt_status cfail(std::string* err, newt_status s, const std::string& msg) {
    if (err) *err = msg;
    return s;
}

// ConvexMesh::validate (geom.rs:318-340).
inline newt_status check_mesh(const double* v, int nv, const int* f, int nf, std::string* err) {
    if (nv < 4) return cfail(err, NEWT_ERR_INVALID_ARG, "a convex mesh needs at least 4 vertices");
    if (nf < 4) return cfail(err, NEWT_ERR_INVALID_ARG, "a convex mesh needs at least 4 triangular faces");
    for (int k = 0; k < 3 * nf; ++k)
        if (f[k] < 0 || f[k] >= nv) return cfail(err, NEWT_ERR_INVALID_ARG, "a mesh face index is out of range");
    for (int k = 0; k < 3 * nv; ++k)
        if (!(v[k] == v[k] && v[k] - v[k] == 0.0))
            return cfail(err, NEWT_ERR_INVALID_ARG, "a mesh vertex has a non-finite coordinate");
    return NEWT_OK;
}

// is_pair_supported (contact.rs:232-257) for the shapes built here.
inline bool pair_supported(int a, int b) {
    auto one = [](int x, int y) {
        return (y == SH_PLANE && (x == SH_SPHERE || x == SH_BOX || x == SH_CAPSULE || x == SH_MESH)) ||
               (x == SH_SPHERE && (y == SH_SPHERE || y == SH_CAPSULE || y == SH_MESH)) ||
               (x == SH_CAPSULE && y == SH_CAPSULE) || (x == SH_BOX && y == SH_BOX) || (x == SH_MESH && y == SH_MESH);
    };
    return one(a, b) || one(b, a);
}

// The mesh tables and the mesh flags of the compiled model.
inline void set_mesh_tables(HostModel* M) {
    ModelT& d = M->dev;
    d.meshes = M->meshes.data();
    d.mesh_v = M->mesh_v.data();
    d.mesh_f = M->mesh_f.data();
    d.n_meshes = (int)M->meshes.size();
    d.has_mesh = 0;
    d.mesh_mesh = 0;
    for (const GeomT& g : M->hgeoms)
        if (g.shape == SH_MESH) d.has_mesh = 1;
    for (const PairT& p : M->pairs)
        if (M->hgeoms[p.a].shape == SH_MESH && M->hgeoms[p.b].shape == SH_MESH) d.mesh_mesh = 1;
}

// Largest contact count one candidate pair can emit (ContactBuf, contact.rs:108-124).
inline int pair_max_contacts(int sa, int sb) {
    if (sa == SH_MESH && sb == SH_MESH) return 4;    // the mesh-mesh manifold
    if (sa == SH_MESH || sb == SH_MESH) return (sa == SH_PLANE || sb == SH_PLANE) ? 2 : 1;  // mesh_plane, sphere_mesh
    if (sa == SH_BOX && sb == SH_BOX) return 4;      // box_box keeps the four deepest
    if (sa != SH_PLANE && sb != SH_PLANE) return 1;  // sphere_vs_sphere_at
    int p = sa == SH_PLANE ? sb : sa;
    if (p == SH_SPHERE) return 1;
    if (p == SH_CAPSULE) return 2;
    return 4;
}

// Parent->child transform of a fixed joint (tree.rs:1124-1131), host double.
inline void fixed_xup(const newt_link_desc& d, double E[9], double r[3]) {
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
inline void spatial_inertia(const newt_link_desc& d, double si[36]) {
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

// Mat3::inverse (math.rs:427-453) of a column-major 3x3, in double; the
// result row-major.  False when the determinant is zero or not finite.
inline bool body_inverse(const double* cm, double* out_rm) {
    auto get = [&](int r, int c) { return cm[c * 3 + r]; };
    double a = get(0, 0), b = get(0, 1), c = get(0, 2);
    double d = get(1, 0), e = get(1, 1), f = get(1, 2);
    double g = get(2, 0), h = get(2, 1), i = get(2, 2);
    double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    if (det == 0.0 || !std::isfinite(det)) return false;
    double inv = 1.0 / det;
    const double col[9] = {(e * i - f * h) * inv, (f * g - d * i) * inv, (d * h - e * g) * inv,
                           (c * h - b * i) * inv, (a * i - c * g) * inv, (b * g - a * h) * inv,
                           (b * f - c * e) * inv, (c * d - a * f) * inv, (a * e - b * d) * inv};
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) out_rm[r * 3 + k] = col[k * 3 + r];
    return true;
}

// A model of free bodies with static geoms (World::add_body, world.rs:950):
// state layout, body tables, geoms riding on bodies, static candidate pairs.
inline newt_status compile_bodies(HostModel* M, std::string* err) {
    const newt_options& o = M->opt;
    if (o.iterations < 1) return cfail(err, NEWT_ERR_INVALID_CONFIG, "solver iterations must be >= 1");
    if (!(o.penetration_slop >= 0.0)) return cfail(err, NEWT_ERR_INVALID_ARG, "penetration_slop must be >= 0");
    if (o.solver_mode == NEWT_SOLVER_NEWTON && o.cone == NEWT_CONE_ELLIPTIC)
        return cfail(err, NEWT_ERR_INVALID_CONFIG, "solver=newton with cone=elliptic is not supported yet; use cone=pyramidal");
    if (o.integrator != NEWT_INTEGRATOR_EULER && o.integrator != NEWT_INTEGRATOR_IMPLICITFAST &&
        o.integrator != NEWT_INTEGRATOR_RK4)
        return cfail(err, NEWT_ERR_INVALID_ARG, "unknown integrator");
    if (o.cone != NEWT_CONE_PYRAMIDAL) return cfail(err, NEWT_ERR_UNSUPPORTED, "elliptic cone is not built");
    const int nb = (int)M->bodies.size();
    M->hbodies.clear();
    M->q_default.clear();
    M->v_default.clear();
    for (int i = 0; i < nb; ++i) {
        const newt_body_desc& b = M->bodies[i];
        if (!(b.mass > 0.0)) return cfail(err, NEWT_ERR_INVALID_ARG, "body mass must be positive");
        BodyT t{};
        t.mass = b.mass;
        t.gravity_scale = b.gravity_scale;
        t.has_max_lin = b.has_max_linear_velocity != 0;
        t.has_max_ang = b.has_max_angular_velocity != 0;
        t.max_lin = b.max_linear_velocity;
        t.max_ang = b.max_angular_velocity;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) t.inertia[r * 3 + c] = b.inertia[c * 3 + r];
        if (!body_inverse(b.inertia, t.inv_inertia))
            return cfail(err, NEWT_ERR_INVALID_ARG, "body inertia tensor must be invertible");
        t.q_off = 7 * i;
        t.v_off = 6 * i;
        M->hbodies.push_back(t);
        for (int k = 0; k < 3; ++k) M->q_default.push_back(b.position[k]);
        for (int k = 0; k < 4; ++k) M->q_default.push_back(b.orientation[k]);
        for (int k = 0; k < 3; ++k) M->v_default.push_back(b.linear_velocity[k]);
        for (int k = 0; k < 3; ++k) M->v_default.push_back(b.angular_velocity_body[k]);
    }
    M->links.clear();
    M->limits.clear();
    M->chain.clear();
    M->hgeoms.clear();
    for (const newt_geom_desc& g : M->geoms) {
        if (g.shape < NEWT_SHAPE_PLANE || g.shape > NEWT_SHAPE_MESH) return cfail(err, NEWT_ERR_INVALID_ARG, "unknown geom shape");
        if (g.shape == NEWT_SHAPE_MESH && (g.mesh < 0 || g.mesh >= (int)M->meshes.size()))
            return cfail(err, NEWT_ERR_INVALID_ARG, "a mesh geom names a mesh that was not added");
        if (g.tree >= 0) return cfail(err, NEWT_ERR_INVALID_ARG, "geom attached to a missing tree");
        if (g.body >= nb) return cfail(err, NEWT_ERR_INVALID_ARG, "geom attached to a missing body");
        if (g.condim != 3 && g.condim != 1 && g.condim != 4 && g.condim != 6)
            return cfail(err, NEWT_ERR_INVALID_ARG, "condim must be 1, 3, 4 or 6");
        GeomT h{};
        h.shape = g.shape;
        h.body = g.body >= 0 ? g.body : -1;
        h.link = h.body;  // pose slot n_links + body with no links
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
        h.mesh = g.shape == NEWT_SHAPE_MESH ? g.mesh : -1;
        M->hgeoms.push_back(h);
    }
    // Static candidate pairs: group/mask and same attachment (Static ==
    // Static, Body(i) == Body(i)) (broadphase.rs:16-18, world.rs:1219-1225).
    M->pairs.clear();
    long cap_c = 0;
    for (size_t a = 0; a < M->geoms.size(); ++a)
        for (size_t b = a + 1; b < M->geoms.size(); ++b) {
            const GeomT& ga = M->hgeoms[a];
            const GeomT& gb = M->hgeoms[b];
            const newt_geom_desc& da = M->geoms[a];
            const newt_geom_desc& db = M->geoms[b];
            if (!((da.group & db.mask) != 0 && (db.group & da.mask) != 0)) continue;
            if (ga.body == gb.body) continue;
            if (!pair_supported(ga.shape, gb.shape))
                return cfail(err, NEWT_ERR_UNSUPPORTED_PAIR, "contact pair geom " + std::to_string(a) + " x geom " +
                                                                 std::to_string(b) + " is an unsupported shape pair");
            if (std::min(ga.condim, gb.condim) != 3)
                return cfail(err, NEWT_ERR_UNSUPPORTED, "only condim 3 pyramidal contacts are built");
            M->pairs.push_back({(int)a, (int)b});
            cap_c += pair_max_contacts(ga.shape, gb.shape);
        }
    // Every contact the pairs can emit fits up to 4096 per world; a world
    // that exceeds a smaller compiled bound faults with NEWT_ERR_CAPACITY.
    const int kMaxBodyContacts = 4096;
    ModelT& d = M->dev;
    d = ModelT{};
    d.n_links = 0;
    d.nq = 7 * nb;
    d.nv = 6 * nb;
    d.n_geoms = (int)M->geoms.size();
    d.n_pairs = (int)M->pairs.size();
    d.cap_c = (int)std::max<long>(std::min<long>(cap_c, kMaxBodyContacts), 1);
    d.cap_r = 4 * d.cap_c;
    d.integrator = o.integrator;
    d.mode = o.solver_mode;
    d.iterations = o.iterations;
    d.cone = o.cone;
    d.dt = o.dt;
    d.slop = o.penetration_slop;
    for (int k = 0; k < 3; ++k) d.gravity[k] = o.gravity[k];
    d.geoms = M->hgeoms.data();
    d.pairs = M->pairs.data();
    d.n_bodies = nb;
    d.bodies = M->hbodies.data();
    d.tnq = d.tnv = 0;
    set_mesh_tables(M);
    if (d.has_mesh && o.solver_mode == NEWT_SOLVER_NEWTON)
        return cfail(err, NEWT_ERR_UNSUPPORTED, "the Newton free-body step runs one wavefront per world and has no "
                                                "mesh contacts; mesh models on free bodies run under PGS or penalty");
    return NEWT_OK;
}

// Sensors (World::add_sensor, world.rs:441; Sensor::validate, sensor.rs:248-323)
// into the device table, with each sensor's first value in sensordata.
// Called after the model's trees or bodies are compiled.
inline newt_status compile_sensors(HostModel* M, std::string* err) {
    ModelT& d = M->dev;
    M->hsensors.clear();
    int out = 0;
    bool dyn = false;
    for (size_t k = 0; k < M->sensors.size(); ++k) {
        const newt_sensor_desc& s = M->sensors[k];
        const std::string who = "sensor " + std::to_string(k) + ": ";
        const int dim = sensor_kind_dim(s.kind);
        if (dim == 0) return cfail(err, NEWT_ERR_UNSUPPORTED, who + "this sensor kind is not built");
        // (tree, link) -> the link's index in the forest
        int gl = -1;
        if (s.tree >= 0 && s.tree < (int)M->trees.size() && s.link >= 0 && s.link < (int)M->trees[s.tree].size()) {
            gl = s.link;
            for (int t = 0; t < s.tree; ++t) gl += (int)M->trees[t].size();
        }
        const bool link_ok = gl >= 0;
        const bool body_ok = s.body >= 0 && s.body < d.n_bodies;
        if (s.kind == SK_JOINT_POS || s.kind == SK_JOINT_VEL) {
            if (!link_ok) return cfail(err, NEWT_ERR_INVALID_CONFIG, who + "link out of range");
            const int j = M->links[gl].joint;
            if (j != JF_HINGE && j != JF_SLIDE)
                return cfail(err, NEWT_ERR_INVALID_CONFIG, who + "joint sensors need a hinge or slide joint");
        } else if (s.kind == SK_FORCE || s.kind == SK_TORQUE || s.kind == SK_SUBTREE_COM) {
            if (!link_ok) return cfail(err, NEWT_ERR_INVALID_CONFIG, who + "link out of range");
        } else if (s.kind == SK_TOUCH) {
            if (s.geom < 0 || s.geom >= d.n_geoms) return cfail(err, NEWT_ERR_INVALID_CONFIG, who + "geom out of range");
        } else if (sensor_kind_site(s.kind)) {
            if (link_ok == body_ok || (s.link >= 0 && !link_ok) || (s.body >= 0 && !body_ok))
                return cfail(err, NEWT_ERR_INVALID_CONFIG, who + "a site rides on exactly one link or free body");
        }
        if (sensor_kind_dynamic(s.kind)) {
            dyn = true;
            if (d.n_bodies > 0)
                return cfail(err, NEWT_ERR_UNSUPPORTED, who + "acceleration, force and touch sensors are built for "
                                                              "models without free bodies only");
            if (d.n_pairs > 0 && d.mode != NEWT_SOLVER_PGS)
                return cfail(err, NEWT_ERR_UNSUPPORTED,
                             who + "acceleration, force and touch sensors with contacts are built for the PGS solver only");
        }
        SensorT t{};
        t.kind = s.kind;
        t.link = link_ok ? gl : -1;
        t.body = s.body;
        t.geom = s.geom;
        t.out = out;
        for (int i = 0; i < 3; ++i) t.off[i] = s.offset[i];
        for (int i = 0; i < 4; ++i) t.quat[i] = s.quat[i];
        M->hsensors.push_back(t);
        out += dim;
    }
    d.n_sensors = (int)M->hsensors.size();
    d.sensor_dim = out;
    d.sensor_dyn = dyn ? 1 : 0;
    for (int i = 0; i < 3; ++i) d.magnetic[i] = M->magnetic[i];
    return NEWT_OK;
}

inline newt_status compile_model(HostModel* M, std::string* err) {
    if (!M->bodies.empty() && M->trees.empty()) return compile_bodies(M, err);
    if (M->trees.empty()) return cfail(err, NEWT_ERR_INVALID_CONFIG, "a model needs a tree or free bodies");
    const int nb = (int)M->bodies.size();  // free bodies beside the trees
    const newt_options& o = M->opt;
    if (o.iterations < 1) return cfail(err, NEWT_ERR_INVALID_CONFIG, "solver iterations must be >= 1");
    if (!(o.penetration_slop >= 0.0)) return cfail(err, NEWT_ERR_INVALID_ARG, "penetration_slop must be >= 0");
    if (o.solver_mode == NEWT_SOLVER_NEWTON && o.cone == NEWT_CONE_ELLIPTIC)
        return cfail(err, NEWT_ERR_INVALID_CONFIG, "solver=newton with cone=elliptic is not supported yet; use cone=pyramidal");
    // ImplicitFast differs from Euler only through actuator velocity terms;
    // models here have no actuators, so both run the Euler step.
    if (o.integrator != NEWT_INTEGRATOR_EULER && o.integrator != NEWT_INTEGRATOR_IMPLICITFAST &&
        o.integrator != NEWT_INTEGRATOR_RK4)
        return cfail(err, NEWT_ERR_INVALID_ARG, "unknown integrator");
    if (o.cone != NEWT_CONE_PYRAMIDAL) return cfail(err, NEWT_ERR_UNSUPPORTED, "elliptic cone is not built");
    // Several trees (World::trees) compile into one forest: tree t's links
    // follow tree t-1's, parents offset by the links before them, so each
    // tree keeps its own root; qpos/qvel concatenate the trees in order.
    std::vector<newt_link_desc> T;
    std::vector<int> link_tree, tree_base;
    for (size_t t = 0; t < M->trees.size(); ++t) {
        tree_base.push_back((int)T.size());
        for (size_t i = 0; i < M->trees[t].size(); ++i) {
            newt_link_desc d = M->trees[t][i];
            if (i == 0 && d.parent >= 0) return cfail(err, NEWT_ERR_INVALID_ARG, "a tree's first link is its root");
            if (i > 0 && !(d.parent >= 0 && d.parent < (int)i))
                return cfail(err, NEWT_ERR_INVALID_ARG, "parent index must precede its link");
            if (i > 0) d.parent += tree_base[t];
            T.push_back(d);
            link_tree.push_back((int)t);
        }
    }
    int nq = 0, nv = 0;
    M->links.clear();
    M->limits.clear();
    M->chain.clear();
    M->q_default.clear();
    for (size_t i = 0; i < T.size(); ++i) {
        const newt_link_desc& d = T[i];
        const bool root = (int)i == tree_base[link_tree[i]];
        if (root && !(d.joint == NEWT_JOINT_FREE || d.joint == NEWT_JOINT_FIXED))
            return cfail(err, NEWT_ERR_INVALID_ARG, "root joint must be Free or Fixed");
        if (!root && d.joint == NEWT_JOINT_FREE) return cfail(err, NEWT_ERR_INVALID_ARG, "free joint only at the root");
        if (d.joint == NEWT_JOINT_BALL) return cfail(err, NEWT_ERR_UNSUPPORTED, "ball joints are not built on the GPU path");
        if (d.has_range && !(d.joint == NEWT_JOINT_HINGE || d.joint == NEWT_JOINT_SLIDE))
            return cfail(err, NEWT_ERR_INVALID_ARG, "a range needs a hinge or slide joint");
        if (d.has_range && o.solver_mode == NEWT_SOLVER_NEWTON)
            return cfail(err, NEWT_ERR_UNSUPPORTED, "Newton joint-limit rows are not built");
        LinkT l{};
        l.lim = -1;
        if (d.has_range) {
            LimitT lt{};
            lt.lo = d.range[0];
            lt.hi = d.range[1];
            // SolRef::DEFAULT (geom.rs:99-102), SolImp::DEFAULT (solver.rs:187-193)
            const double sr[2] = {0.02, 1.0}, sim[4] = {0.9, 0.95, 0.001, 0.5};
            for (int k = 0; k < 2; ++k) lt.solref[k] = d.limit_custom ? d.limit_solref[k] : sr[k];
            for (int k = 0; k < 4; ++k) lt.solimp[k] = d.limit_custom ? d.limit_solimp[k] : sim[k];
            lt.power = d.limit_custom ? d.limit_solimp_power : 2;
            // JointLimit::DEFAULT stiffness / damping (joint.rs:152-157)
            lt.stiffness = d.limit_custom ? d.limit_stiffness : 1000.0;
            lt.damping = d.limit_custom ? d.limit_damping : 60.0;
            lt.link = (int)i;
            l.lim = (int)M->limits.size();
            M->limits.push_back(lt);
        }
        l.parent = root ? -1 : d.parent;
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
        for (int j = (int)i; j >= 0; j = j == tree_base[link_tree[j]] ? -1 : T[j].parent) ch.push_back(j);
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
    // level schedule and sibling ranks
    {
        int n = (int)M->links.size();
        std::vector<int> nchild(n, 0);
        int maxd = 0;
        for (int i = 0; i < n; ++i) {
            LinkT& l = M->links[i];
            l.depth = l.parent < 0 ? 0 : M->links[l.parent].depth + 1;
            l.rank = l.parent < 0 ? 0 : nchild[l.parent]++;
            for (int w = 0; w < 2; ++w) l.anc[w] = l.parent < 0 ? 0ull : M->links[l.parent].anc[w];
            if (i < 128) l.anc[i >> 6] |= 1ull << (i & 63);
            maxd = std::max(maxd, l.depth);
        }
        int max_rank = 1;
        for (int c : nchild) max_rank = std::max(max_rank, c);
        M->level_off.assign(1, 0);
        M->level_links.clear();
        for (int dl = 0; dl <= maxd; ++dl) {
            for (int i = 0; i < n; ++i)
                if (M->links[i].depth == dl) M->level_links.push_back(i);
            M->level_off.push_back((int)M->level_links.size());
        }
        M->dev.n_levels = maxd + 1;
        M->dev.max_rank = max_rank;
        M->level_nrank.assign(maxd + 1, 0);
        M->dof_link.clear();
        for (int i = 0; i < n; ++i) {
            const LinkT& l = M->links[i];
            int cnt = l.joint == NEWT_JOINT_FREE ? 6 : (l.joint == NEWT_JOINT_FIXED ? 0 : 1);
            for (int k = 0; k < cnt; ++k) M->dof_link.push_back(i);
        }
        M->dev.damped = 0;
        for (int i = 0; i < n; ++i) {
            const LinkT& l = M->links[i];
            M->level_nrank[l.depth] = std::max(M->level_nrank[l.depth], l.rank + 1);
            // RK4 solves contacts with M itself (no implicit damping term).
            if (o.integrator != NEWT_INTEGRATOR_RK4 &&
                ((l.joint == NEWT_JOINT_FREE && l.free_damping != 0.0) ||
                 ((l.joint == NEWT_JOINT_HINGE || l.joint == NEWT_JOINT_SLIDE) && l.damping != 0.0)))
                M->dev.damped = 1;
        }
    }
    // geoms
    M->hgeoms.clear();
    for (const newt_geom_desc& g : M->geoms) {
        if (g.shape < NEWT_SHAPE_PLANE || g.shape > NEWT_SHAPE_MESH) return cfail(err, NEWT_ERR_INVALID_ARG, "unknown geom shape");
        if (g.shape == NEWT_SHAPE_MESH && (g.mesh < 0 || g.mesh >= (int)M->meshes.size()))
            return cfail(err, NEWT_ERR_INVALID_ARG, "a mesh geom names a mesh that was not added");
        if (g.tree >= (int)M->trees.size()) return cfail(err, NEWT_ERR_INVALID_ARG, "geom attached to a missing tree");
        if (g.tree >= 0 && (g.link < 0 || g.link >= (int)M->trees[g.tree].size()))
            return cfail(err, NEWT_ERR_INVALID_ARG, "geom link out of range");
        if (g.body >= nb) return cfail(err, NEWT_ERR_INVALID_ARG, "geom attached to a missing body");
        if (g.tree >= 0 && g.body >= 0) return cfail(err, NEWT_ERR_INVALID_ARG, "a geom rides on a link or a body");
        if (g.condim != 3 && g.condim != 1 && g.condim != 4 && g.condim != 6)
            return cfail(err, NEWT_ERR_INVALID_ARG, "condim must be 1, 3, 4 or 6");
        GeomT h{};
        h.shape = g.shape;
        // A body's geoms take the pose slot after the links.
        h.link = g.tree >= 0 ? tree_base[g.tree] + g.link : g.body >= 0 ? (int)T.size() + g.body : -1;
        h.body = g.body >= 0 ? g.body : -1;
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
        h.mesh = g.shape == NEWT_SHAPE_MESH ? g.mesh : -1;
        M->hgeoms.push_back(h);
    }
    // Static candidate pairs: group/mask filter, disabled self collision,
    // same attachment (broadphase.rs:16-18, 426-437; world.rs:1219-1225).
    M->pairs.clear();
    int cap_c = 0;
    for (size_t a = 0; a < M->geoms.size(); ++a)
        for (size_t b = a + 1; b < M->geoms.size(); ++b) {
            const newt_geom_desc& ga = M->geoms[a];
            const newt_geom_desc& gb = M->geoms[b];
            if (!((ga.group & gb.mask) != 0 && (gb.group & ga.mask) != 0)) continue;
            bool la = ga.tree >= 0, lb = gb.tree >= 0;
            const bool ba = ga.body >= 0, bb = gb.body >= 0;
            // disabled_self_collision holds a tree whose links skip each other
            if (la && lb && ga.tree == gb.tree && M->tree_collide[ga.tree] == 0) continue;
            // the same attachment: Static, Body(i) or Link(t, l)
            if (!la && !lb && !ba && !bb) continue;
            if (ba && bb && ga.body == gb.body) continue;
            if (la && lb && ga.tree == gb.tree && ga.link == gb.link) continue;
            // newt's penalty step runs the bodies against the trees' states
            // of that moment (world.rs:1705-1785, 1826-1946): not built.
            if (nb > 0 && o.solver_mode == NEWT_SOLVER_PENALTY)
                return cfail(err, NEWT_ERR_UNSUPPORTED, "penalty contacts in a model with trees and free bodies are "
                                                        "not built (geom " + std::to_string(a) + " x geom " +
                                                            std::to_string(b) + ")");
            // newt steps trees one after another under the penalty solver
            // (step_trees, world.rs:1826-1946), each against the others'
            // states of that moment: not built.
            if (la && lb && ga.tree != gb.tree && o.solver_mode == NEWT_SOLVER_PENALTY)
                return cfail(err, NEWT_ERR_UNSUPPORTED,
                             "penalty contacts between two trees are not built (geom " + std::to_string(a) +
                                 " x geom " + std::to_string(b) + ")");
            const std::string which = "contact pair geom " + std::to_string(a) + " x geom " + std::to_string(b);
            if (!pair_supported(ga.shape, gb.shape))
                return cfail(err, NEWT_ERR_UNSUPPORTED_PAIR, which + " is an unsupported shape pair");
            if (std::min(ga.condim, gb.condim) != 3)
                return cfail(err, NEWT_ERR_UNSUPPORTED, "only condim 3 pyramidal contacts are built");
            M->pairs.push_back({(int)a, (int)b});
            cap_c += pair_max_contacts(ga.shape, gb.shape);
        }
    M->v_default.assign(nv, 0.0);
    // Free bodies follow the trees in qpos / qvel (World state order).
    M->hbodies.clear();
    for (int i = 0; i < nb; ++i) {
        const newt_body_desc& b = M->bodies[i];
        if (!(b.mass > 0.0)) return cfail(err, NEWT_ERR_INVALID_ARG, "body mass must be positive");
        BodyT t{};
        t.mass = b.mass;
        t.gravity_scale = b.gravity_scale;
        t.has_max_lin = b.has_max_linear_velocity != 0;
        t.has_max_ang = b.has_max_angular_velocity != 0;
        t.max_lin = b.max_linear_velocity;
        t.max_ang = b.max_angular_velocity;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) t.inertia[r * 3 + c] = b.inertia[c * 3 + r];
        if (!body_inverse(b.inertia, t.inv_inertia))
            return cfail(err, NEWT_ERR_INVALID_ARG, "body inertia tensor must be invertible");
        t.q_off = nq + 7 * i;
        t.v_off = nv + 6 * i;
        M->hbodies.push_back(t);
        for (int k = 0; k < 3; ++k) M->q_default.push_back(b.position[k]);
        for (int k = 0; k < 4; ++k) M->q_default.push_back(b.orientation[k]);
        for (int k = 0; k < 3; ++k) M->v_default.push_back(b.linear_velocity[k]);
        for (int k = 0; k < 3; ++k) M->v_default.push_back(b.angular_velocity_body[k]);
    }
    ModelT& d = M->dev;
    d.n_links = (int)T.size();
    d.nq = nq + 7 * nb;
    d.nv = nv + 6 * nb;
    d.tnq = nq;
    d.tnv = nv;
    d.forest = M->trees.size() > 1 ? 1 : 0;
    d.n_bodies = nb;
    d.bodies = nb > 0 ? M->hbodies.data() : nullptr;
    // Penalty limits and penalty contacts run on the lane-per-world FP64 step.
    d.penalty_limits = (!M->limits.empty() || !M->pairs.empty()) && o.solver_mode == NEWT_SOLVER_PENALTY;
    d.n_geoms = (int)M->geoms.size();
    d.n_pairs = (int)M->pairs.size();
    d.cap_c = std::max(cap_c, 1);
    d.cap_r = std::max(4 * cap_c, 1);
    d.n_chain = (int)M->chain.size();
    d.integrator = o.integrator;
    d.n_ranged = (int)M->limits.size();
    d.ext = d.n_ranged > 0 || o.integrator == NEWT_INTEGRATOR_RK4 ? 1 : 0;
    for (const PairT& pr : M->pairs)
        if (M->hgeoms[pr.a].shape != SH_PLANE && M->hgeoms[pr.b].shape != SH_PLANE) d.ext = 1;
    d.mode = o.solver_mode;
    d.iterations = o.iterations;
    d.cone = o.cone;
    d.dt = o.dt;
    d.slop = o.penetration_slop;
    for (int k = 0; k < 3; ++k) d.gravity[k] = o.gravity[k];
    d.links = M->links.data();
    d.geoms = M->hgeoms.data();
    d.pairs = M->pairs.data();
    d.chain = M->chain.data();
    d.level_off = M->level_off.data();
    d.level_links = M->level_links.data();
    d.level_nrank = M->level_nrank.data();
    d.dof_link = M->dof_link.data();
    set_mesh_tables(M);
    return NEWT_OK;
}

// Tree models the persistent wavefront step does not run: ranged joints
// under the penalty solver, several trees, the Newton contact solve and
// convex meshes step on the lane-per-world FP64 kernel in every precision.
inline bool lane_only(const ModelT& m) {
    return (m.n_bodies == 0 && (m.penalty_limits || m.forest || m.mode == NEWT_SOLVER_NEWTON || m.has_mesh)) ||
           (m.n_bodies > 0 && m.n_links > 0);
}

// Per-world mesh-mesh workspace (newt_ccd.hpp, Work): the EPA polytope
// (128 vertices, two buffers of 512 faces, 1024 horizon edges) and the
// manifold's five 32-point polygons; only when some pair is mesh-mesh.
template <class F> void add_ccd_layout(const ModelT& m, Layout& L, F D) {
    const bool on = m.mesh_mesh != 0;
    L.cv = on ? D(9 * 128) : 0;
    L.cf0 = on ? D(4 * 512) : 0;
    L.cf1 = on ? D(4 * 512) : 0;
    L.cpoly = on ? D(5 * 96) : 0;
}
template <class F> void add_ccd_layout_i(const ModelT& m, Layout& L, F I) {
    const bool on = m.mesh_mesh != 0;
    L.cvi = on ? I(2 * 128) : 0;
    L.cfi0 = on ? I(3 * 512) : 0;
    L.cfi1 = on ? I(3 * 512) : 0;
    L.ce = on ? I(2 * 1024) : 0;
}

inline Layout make_layout(const ModelT& m, int real_bytes) {
    Layout L{};
    int r = 0, dd = 0, ii = 0;
    auto R = [&](int n) { int o = r; r += n; return o; };
    auto D = [&](int n) { int o = dd; dd += n; return o; };
    auto I = [&](int n) { int o = ii; ii += n; return o; };
    // The tree part of the state; free bodies beside the trees add their
    // pose slots and the free-body solve's workspace.
    int n = m.n_links, nv = std::max(m.tnv, 1), nb = m.n_bodies;
    L.coef = R(2 * nv * m.cap_r);
    L.A = R(m.cap_r * m.cap_r);
    L.total_r = r;
    L.poses = D(7 * (n + nb));
    L.lvel = D(6 * n);
    L.xup = D(12 * n);
    L.S = D(6 * n);
    L.v = D(6 * n);
    L.c = D(6 * n);
    L.a = D(6 * n);
    L.ia = D(36 * n);
    L.pa = D(6 * n);
    L.d = D(n);
    L.fs = D(n);
    L.tau = D(nv);
    L.qacc_free = D(nv);
    L.delta = D(nv);
    L.gres = D(nv);
    L.jt = D(3 * nv);
    L.cgeo = D(9 * m.cap_c);
    L.iwc = D(n);
    L.rreg = D(m.cap_r);
    L.rdiag = D(m.cap_r);
    // row biases; under penalty contacts the per-link contact wrenches instead
    L.rbias = D(std::max(m.cap_r, m.mode == NEWT_SOLVER_PENALTY ? 6 * m.n_links : 0));
    L.f = D(m.cap_r);
    L.M = D(nv * nv);
    L.Li = D(nv * nv);
    L.Ld = D(nv * nv);
    L.dvec = D(nv);
    L.dres = D(nv);
    // RK4 (tree.rs:1234-1364): start state, stage derivative, weighted sums.
    const bool rk4 = m.integrator == NEWT_INTEGRATOR_RK4;
    L.rq0 = rk4 ? D(m.tnq) : 0;
    L.rv0 = rk4 ? D(nv) : 0;
    L.rdq = rk4 ? D(m.tnq) : 0;
    L.raq = rk4 ? D(m.tnq) : 0;
    L.rav = rk4 ? D(nv) : 0;
    const int nr = m.n_ranged;
    L.lim = D(nv);
    L.ldv = D(std::max(nr * nv, 1));
    L.lA = D(std::max(nr * nr, 1));
    L.lb = D(std::max(nr, 1));
    L.ldg = D(std::max(nr, 1));
    L.lf = D(std::max(nr, 1));
    L.lc = D(std::max(nr, 1));
    L.lvi = D(std::max(nr, 1));
    const bool sens = m.n_sensors > 0;
    L.swr = sens ? D(6 * n) : 0;
    L.sva = sens ? D(12 * n) : 0;
    L.sfl = sens ? D(6 * n) : 0;
    L.sqdd = sens ? D(nv) : 0;
    L.sfn = sens ? D(std::max(m.cap_c, 1)) : 0;
    const bool newton = m.mode == NEWT_SOLVER_NEWTON;
    L.nfac = newton ? D(m.cap_r * m.cap_r) : 0;
    L.nvec = newton ? D(9 * m.cap_r) : 0;
    L.nbrk = newton ? D(m.cap_r + 2) : 0;
    if (nb > 0) {
        L.bwr = D(6 * nb);
        L.biw = D(9 * nb);
        L.bbd = D(6 * nb);
        L.brow = D(7 * m.cap_r);
        L.bblk = D(8 * m.cap_c);
        L.bcomp = D(12 * m.cap_r);
        L.bdel = D(6);
        L.tbw = D(6 * nb);
    }
    add_ccd_layout(m, L, D);
    L.total_d = dd;
    L.cids = I(3 * m.cap_c);
    L.rinfo = I(3 * m.cap_r);
    L.pairs = I(std::max(m.n_pairs, 1));
    L.iwf = I(n + 2);
    L.lsl = I(std::max(m.n_ranged, 1));
    L.llm = I(std::max(m.n_ranged, 1));
    L.bbi = nb > 0 ? I(2 * m.cap_c) : 0;
    add_ccd_layout_i(m, L, I);
    L.total_i = ii;
    (void)real_bytes;
    return L;
}


}  // namespace newt

namespace newt {

// Per-world workspace of a free-body model: body poses, candidate pairs,
// contact geometry and the free-body solve (rows 4 per pyramidal contact).
inline Layout make_body_layout(const ModelT& m) {
    Layout L{};
    int dd = 0, ii = 0;
    auto D = [&](int n) { int o = dd; dd += n; return o; };
    auto I = [&](int n) { int o = ii; ii += n; return o; };
    const int nb = m.n_bodies;
    L.poses = D(7 * nb);
    L.cgeo = D(9 * m.cap_c);
    L.bwr = D(6 * nb);
    L.biw = D(9 * nb);
    L.bbd = D(6 * nb);
    // rows; under the penalty solver the RK4 stage scratch (s0 and the
    // running stage sum, 26 per body) instead
    L.brow = D(std::max(7 * m.cap_r, m.mode == NEWT_SOLVER_PENALTY ? 26 * nb : 0));
    L.bblk = D(8 * m.cap_c);
    add_ccd_layout(m, L, D);
    L.total_d = std::max(dd, 1);
    L.cids = I(3 * m.cap_c);
    L.pairs = I(std::max(m.n_pairs, 1));
    L.iwf = I(2);
    L.bbi = I(2 * m.cap_c);
    add_ccd_layout_i(m, L, I);
    L.total_i = ii;
    return L;
}

}  // namespace newt
