// newt_types.hpp — device tables of a compiled model and the per-world
// workspace layout.
//
// A model is compiled once on the host (newt_model_finalize) into flat
// arrays that every world of a batch shares: link records, geom records, the
// statically filtered broadphase candidate pairs and each link's ancestor
// chain.  Per-world state and scratch live in "lane-interleaved" arrays:
// element i of world w sits at base[i * W + w], so the lanes of a wavefront,
// which step consecutive worlds, touch consecutive addresses.
#pragma once

#include <stdint.h>

namespace newt {

enum : int { JF_FREE = 0, JF_FIXED = 1, JF_HINGE = 2, JF_SLIDE = 3, JF_BALL = 4 };
enum : int { SH_PLANE = 0, SH_SPHERE = 1, SH_BOX = 2, SH_CAPSULE = 3, SH_MESH = 4 };

struct LinkT {
    int parent, joint, q_off, v_off;
    int chain_off, chain_len;  // ancestors root..self in ModelT::chain
    int depth, rank;           // tree depth; index among the parent's children
    unsigned long long anc[2]; // bit j set when link j is this link or an ancestor (links < 128)
    double axis[3];            // as given (joint subspace)
    double pj[3], cj[3];       // joint anchor in parent / child frame
    double pjq[4], cjq[4];     // anchor orientations (fixed joints, root pose)
    double xfE[9], xfr[3];     // fixed-joint parent->child transform
    double si[36];             // 6x6 spatial inertia at the COM
    double mass, damping, armature, free_damping;
    int lim, pad1;             // index into ModelT::limits, -1 without a range
};

// Range of a hinge/slide with its PGS limit-row parameters (JointKind range,
// JointLimit::solref/solimp, joint.rs:124-141).
struct LimitT {
    double lo, hi;
    double solref[2], solimp[4];
    int power, link;
    double stiffness, damping;  // penalty spring (penalty solver)
};

struct GeomT {
    int shape, link;           // pose slot: the link, or n_links + body; -1 for a static geom
    int condim, max_contacts;
    double size[3], pos[3], quat[4];
    double friction, margin, gap;
    double solref[2], solimp[4];
    int power, body;           // body = the free body it rides on, -1 for none
    int mesh, pad;             // SH_MESH: index into ModelT::meshes
};

// A convex mesh (ConvexMesh, geom.rs:305-312): vertices mesh_v[3 v_off ..
// 3 (v_off + nv)) (x, y, z), faces mesh_f[3 f_off .. 3 (f_off + nf)).
struct MeshT {
    int v_off, nv, f_off, nf;
};

// A free Body (body.rs:22-47).  Inertia and its inverse row-major m[r*3+c];
// its state is qpos[q_off..+7) = (position, orientation x y z w) and
// qvel[v_off..+6) = (linear velocity in world, angular velocity in body).
struct BodyT {
    double mass, gravity_scale, max_lin, max_ang;
    double inertia[9], inv_inertia[9];
    int has_max_lin, has_max_ang, q_off, v_off;
};

struct PairT {
    int a, b;
};

enum : int {
    SK_JOINT_POS = 0, SK_JOINT_VEL = 1, SK_FRAME_POS = 4, SK_FRAME_QUAT = 5, SK_GYRO = 6, SK_ACCEL = 7,
    SK_TOUCH = 8, SK_FORCE = 9, SK_TORQUE = 10, SK_VELOCIMETER = 13, SK_MAGNETOMETER = 14, SK_SUBTREE_COM = 16,
    SK_FRAME_LINVEL = 17, SK_FRAME_ANGVEL = 18
};

// SensorKind::dim (sensor.rs:200-221); 0 for kinds this build does not
// evaluate.
inline __host__ __device__ int sensor_kind_dim(int kind) {
    switch (kind) {
        case SK_JOINT_POS: case SK_JOINT_VEL: case SK_TOUCH: return 1;
        case SK_FRAME_QUAT: return 4;
        case SK_FRAME_POS: case SK_GYRO: case SK_ACCEL: case SK_FORCE: case SK_TORQUE: case SK_VELOCIMETER:
        case SK_MAGNETOMETER: case SK_SUBTREE_COM: case SK_FRAME_LINVEL: case SK_FRAME_ANGVEL: return 3;
        default: return 0;
    }
}
// Kinds that read the post-step accelerations or contact forces.
inline __host__ __device__ bool sensor_kind_dynamic(int kind) {
    return kind == SK_ACCEL || kind == SK_FORCE || kind == SK_TORQUE || kind == SK_TOUCH;
}
inline __host__ __device__ bool sensor_kind_site(int kind) {
    return kind == SK_FRAME_POS || kind == SK_FRAME_QUAT || kind == SK_GYRO || kind == SK_ACCEL ||
           kind == SK_VELOCIMETER || kind == SK_MAGNETOMETER || kind == SK_FRAME_LINVEL || kind == SK_FRAME_ANGVEL;
}
// Kinds that read the links' spatial velocities (sensor.rs:463-474).
inline __host__ __device__ bool sensor_kind_va(int kind) {
    return kind == SK_GYRO || kind == SK_ACCEL || kind == SK_FORCE || kind == SK_TORQUE || kind == SK_VELOCIMETER ||
           kind == SK_FRAME_LINVEL || kind == SK_FRAME_ANGVEL;
}

// A sensor (sensor.rs:226-229) and its site frame (sensor.rs:118-122):
// `out` is its first value in a world's sensordata.
struct SensorT {
    int kind, link, body, geom, out, pad;
    double off[3], quat[4];
};

struct ModelT {
    int n_links, nq, nv, n_geoms;
    int n_pairs, cap_c, cap_r, n_chain;
    int integrator, mode, iterations, cone;
    double dt, slop;
    double gravity[3];
    const LinkT* links;
    const GeomT* geoms;
    const PairT* pairs;
    const int* chain;
    // Level schedule: links of depth d are level_links[level_off[d] ..
    // level_off[d+1]); max_rank = largest sibling count; level_nrank[d] =
    // largest sibling count among the links of depth d.
    int n_levels, max_rank;
    const int* level_off;
    const int* level_links;
    const int* level_nrank;
    const int* dof_link;  // [nv] link owning each dof
    int damped;    // some joint has damping or free_damping != 0 (Euler)
    int n_ranged;  // hinges/slides with a range
    const LimitT* limits;
    int ext;       // needs the extended persistent step: limits, RK4 or non-plane pairs
    int n_bodies;  // free bodies; with trees too, their state follows the trees' (q_off / v_off)
    int tnq, tnv;  // the trees' share of nq / nv (all of it without free bodies, 0 without trees)
    int penalty_limits;  // ranged joints under the penalty solver: the lane-per-world step only
    const BodyT* bodies;
    // Sensors, evaluated after the last step of a launch (sensor.rs:446-702).
    const SensorT* sensors;
    int n_sensors, sensor_dim;
    int sensor_dyn;  // some sensor needs the post-step accelerations or contact forces
    int forest;      // several trees (several roots): the lane-per-world step only
    double magnetic[3];  // World::magnetic_field
    // Convex meshes (World::meshes) and whether some geom is a mesh (the
    // lane-per-world steps only) or some pair is mesh-mesh (EPA workspace).
    const MeshT* meshes;
    const double* mesh_v;
    const int* mesh_f;
    int n_meshes, has_mesh, mesh_mesh;
};

// Offsets (in elements) of the per-world scratch regions.
struct Layout {
    // Jacobian rows and Delassus matrix (type R of the step: FP64 or FP32)
    int coef, A;
    int total_r;
    // FP64 workspace: kinematics, dynamics, contact geometry, factorization
    int poses, lvel, xup, S, v, c, a, ia, pa, d, fs, tau, qacc_free, delta, gres;
    int jt, cgeo, iwc, rreg, rdiag, rbias, f, M, Li, Ld, dvec, dres;
    int rq0, rv0, rdq, raq, rav;  // RK4 only
    // joint-limit rows (up to n_ranged): force delta [nv], M⁻¹ e_i, A, and
    // per-row bias, diagonal, impulse, coefficient, violation
    int lim, ldv, lA, lb, ldg, lf, lc, lvi;
    // free-body solve: per body the wrench (6), world inverse inertia (9)
    // and impulse accumulator (6); per row dir (3), reg, diag, bias,
    // impulse; per contact block the two arms (6) and the two facet-pair
    // cross responses
    int bwr, biw, bbd, brow, bblk;
    // trees with free bodies: body components of the tree rows (linear and
    // angular Jacobian, 6 per row and side), one column's body response (6)
    // and the tree solve's body wrenches (6 per body)
    int bcomp, bdel, tbw;
    // sensors: per-link contact wrenches (6), v and a (12) and joint
    // wrenches (6); qddot [nv]; per-contact normal forces
    int swr, sva, sfl, sqdd, sfn;
    // Newton contact solve on a tree: the factor of H = A + diag(reg)
    // [cap_r^2], vectors [9 cap_r] and line-search breakpoints [cap_r + 2]
    int nfac, nvec, nbrk;
    // mesh-mesh contacts: EPA vertices (9 each, 128), two face buffers (4
    // each, 512), the manifold's polygons (5 x 32 points)
    int cv, cf0, cf1, cpoly;
    int total_d;
    // integer workspace
    int cids, rinfo, pairs, iwf;
    int lsl, llm;  // limit rows: dof slot, limit index
    int bbi;       // free-body contact blocks: body_a, body_b
    int cvi, cfi0, cfi1, ce;  // EPA tie flags (2 each), face indices (3 each), horizon edges (2 each, 1024)
    int total_i;
};

struct DataT {
    int W;            // worlds in this batch
    double* qpos;     // [nq][W]
    double* qvel;     // [nv][W]
    double* qacc;     // [nv][W]
    double* qfrc;     // [nv][W] user qfrc_applied
    double* fat;      // [n_geoms*6][W] broadphase fat bounds
    int* fault;       // [W]
    int* ncontact;    // [W] contacts of the last step
    int* npair;       // [W] pairs of the last step
    // Nonzero: rebuild every proxy's fat bound from the state before the
    // first step (the refresh the state setters make, world.rs:1119-1136)
    // instead of continuing from `fat`.  Read by the persistent step.
    int init_bounds;
    double* sensordata;  // [sensor_dim][W]
};

}  // namespace newt
