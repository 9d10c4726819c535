/*
 * newt_gpu.h — C ABI of the MI355X batched newt step.
 *
 * The reference engine has one mutable `World` (model and state together,
 * newt/src/world.rs:108-204) advanced by `World::step(&mut self)`
 * (newt/src/world.rs:1459-1537).  This boundary splits it into an immutable
 * `newt_model` (the non-state fields of `World`: dt, gravity, integrator,
 * solver config, trees' links, geoms, pair filters) and a `newt_data`
 * holding the mutable state of N independent worlds that share that model.
 * `newt_step` advances every world of a `newt_data` by `n_steps` calls of
 * the reference's `World::step`.
 *
 * All functions return a `newt_status`; on failure `newt_last_error()`
 * returns a thread-local message.  The reference panics instead
 * (newt/src/solver.rs:123-131, newt/src/world.rs:874-898,
 * newt/src/solver.rs:2856-2858); here a panic-class condition inside a
 * world sets that world's fault flag, freezes it, and the other worlds
 * continue.
 *
 * Plain C types only: pointers, sizes, doubles.  Host arrays are row-major
 * [world][slot].
 */
#ifndef NEWT_GPU_H
#define NEWT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum newt_status {
    NEWT_OK = 0,
    NEWT_ERR_INVALID_CONFIG = 1,  /* SolverConfig::validate (solver.rs:123-131) */
    NEWT_ERR_UNSUPPORTED_PAIR = 2,/* World::assert_pairs_supported (world.rs:874-898) */
    NEWT_ERR_NOT_PD = 3,          /* tree contact solve, non-PD M (solver.rs:2856) */
    NEWT_ERR_NEWTON_FAIL = 4,     /* Newton solve failure (solver.rs:1379) */
    NEWT_ERR_OOM = 5,
    NEWT_ERR_DEVICE = 6,
    NEWT_ERR_INVALID_ARG = 7,
    NEWT_ERR_CAPACITY = 8,        /* a world exceeded a compiled capacity */
    NEWT_ERR_UNSUPPORTED = 9      /* a model feature this build does not step */
} newt_status;

/* JointKind (newt/src/joint.rs:40-110). */
typedef enum newt_joint_kind {
    NEWT_JOINT_FREE = 0,
    NEWT_JOINT_FIXED = 1,
    NEWT_JOINT_HINGE = 2,
    NEWT_JOINT_SLIDE = 3,
    NEWT_JOINT_BALL = 4
} newt_joint_kind;

/* GeomShape (newt/src/geom.rs:189). */
typedef enum newt_shape_kind {
    NEWT_SHAPE_PLANE = 0,
    NEWT_SHAPE_SPHERE = 1,
    NEWT_SHAPE_BOX = 2,
    NEWT_SHAPE_CAPSULE = 3,
    NEWT_SHAPE_MESH = 4           /* GeomShape::Mesh (geom.rs:220): a convex mesh */
} newt_shape_kind;

/* Integrator (newt/src/world.rs:76-94). */
typedef enum newt_integrator {
    NEWT_INTEGRATOR_EULER = 0,
    NEWT_INTEGRATOR_IMPLICITFAST = 1,
    NEWT_INTEGRATOR_RK4 = 2
} newt_integrator;

/* SolverMode (newt/src/solver.rs:62-73). */
typedef enum newt_solver_mode {
    NEWT_SOLVER_PENALTY = 0,
    NEWT_SOLVER_PGS = 1,
    NEWT_SOLVER_NEWTON = 2
} newt_solver_mode;

/* ConeKind (newt/src/solver.rs:77-89). */
typedef enum newt_cone {
    NEWT_CONE_PYRAMIDAL = 0,
    NEWT_CONE_ELLIPTIC = 1
} newt_cone;

/* Arithmetic of a newt_data. */
typedef enum newt_precision {
    /* Every quantity in FP64; the drop-in for World::step.  The contact
     * solve runs one wavefront per world (parallel Delassus assembly,
     * incremental-residual PGS); results agree with the CPU step to ~1e-12. */
    NEWT_PRECISION_FP64 = 0,
    /* FP64 kinematics, dynamics (CRB, ABA), Cholesky, residuals, impulses and
     * qpos/quaternion accumulation; Jacobian rows and the Delassus matrix
     * are stored in FP32.  qacc within 1e-5 of the FP64 step. */
    NEWT_PRECISION_MIXED = 1,
    /* FP64 with newt's operation order everywhere, one lane per world:
     * bit-identical to the CPU restatement of World::step.  For
     * certification; slow when contact sets are large. */
    NEWT_PRECISION_FP64_EXACT = 2
} newt_precision;

/* World options: the scalar, non-state fields of World
 * (newt/src/world.rs:110-166). */
typedef struct newt_options {
    double dt;               /* World::dt */
    double gravity[3];       /* World::gravity */
    double penetration_slop; /* World::penetration_slop (>= 0): raises every
                                contact's force-free gap to at least itself
                                (append_contacts, world.rs:2517-2528) */
    int32_t integrator;      /* newt_integrator */
    int32_t solver_mode;     /* newt_solver_mode */
    int32_t iterations;      /* SolverConfig::iterations */
    int32_t cone;            /* newt_cone */
} newt_options;

/* One Link (newt/src/tree.rs:79-123) plus its joint (joint.rs:40-110).
 * Quaternions are (x, y, z, w); the inertia is column-major like Mat3
 * (newt/src/math.rs:386). */
typedef struct newt_link_desc {
    int32_t parent;          /* -1 for the root */
    int32_t joint;           /* newt_joint_kind */
    double axis[3];          /* hinge/slide axis in the parent frame */
    double damping;          /* JointKind damping */
    double armature;         /* JointKind armature */
    double free_damping;     /* Link::free_damping */
    int32_t has_range;       /* JointKind range Some(..) */
    int32_t pad0;
    double range[2];
    double parent_pos[3];    /* joint_offset_in_parent.0 */
    double parent_quat[4];   /* joint_offset_in_parent.1 */
    double child_pos[3];     /* joint_offset_in_child.0 */
    double child_quat[4];    /* joint_offset_in_child.1 */
    double mass;
    double inertia[9];       /* inertia_body, column-major */
    /* JointLimit (joint.rs:124-165): solref / solimp of the PGS limit row
     * and the penalty spring (stiffness, damping) of the penalty solver;
     * limit_custom = 0 takes JointLimit::DEFAULT (1000, 60, SolRef::DEFAULT,
     * SolImp::DEFAULT). */
    int32_t limit_custom;
    int32_t limit_solimp_power;
    double limit_solref[2];
    double limit_solimp[4];
    double limit_stiffness;
    double limit_damping;
} newt_link_desc;

/* One free Body (newt/src/body.rs:22-47): inertia about the COM in body
 * axes (column-major, inverted on the host as Body::new does, body.rs:53-70),
 * the initial state every world starts from, the optional speed caps and
 * the gravity scale.  rolling_friction is not modelled (None). */
typedef struct newt_body_desc {
    double mass;
    double inertia[9];
    double position[3];              /* world COM */
    double orientation[4];           /* body -> world (x, y, z, w) */
    double linear_velocity[3];       /* world frame */
    double angular_velocity_body[3]; /* body frame */
    int32_t has_max_linear_velocity;
    int32_t has_max_angular_velocity;
    double max_linear_velocity;
    double max_angular_velocity;
    double gravity_scale;            /* 1.0 in Body::new */
} newt_body_desc;

/* One Geom (newt/src/geom.rs:352-419). */
typedef struct newt_geom_desc {
    int32_t shape;           /* newt_shape_kind */
    int32_t tree;            /* attached tree index, -1 for none */
    int32_t link;            /* attached link index within that tree */
    int32_t condim;          /* 1, 3, 4 or 6 */
    double size[3];          /* sphere: r; box: half extents; capsule: r, half_height */
    double pos[3];           /* local_offset */
    double quat[4];          /* local_orientation (x, y, z, w) */
    double friction;
    double margin;
    double gap;
    double solref[2];        /* (timeconst, dampratio) */
    double solimp[4];        /* (dmin, dmax, width, midpoint) */
    int32_t solimp_power;    /* SolImp::power */
    uint32_t group;          /* collision_group */
    uint32_t mask;           /* collision_mask */
    int32_t body;            /* Geom::body: attached free body, -1 for none */
    int32_t mesh;            /* NEWT_SHAPE_MESH: the newt_model_add_mesh index */
} newt_geom_desc;

typedef struct newt_model newt_model;
typedef struct newt_data newt_data;

/* Thread-local description of the last failure. */
const char* newt_last_error(void);

/* ---- model: the immutable half of World ------------------------------ */

/* World::new + the option fields (newt/src/world.rs:326-361). */
newt_status newt_model_create(const newt_options* options, newt_model** out);

/* World::add_tree (newt/src/world.rs:965).  `self_collide` = 0 adds the
 * tree to World::disabled_self_collision (newt/src/world.rs:130).  Several
 * trees share one world: their coordinates follow each other in qpos/qvel
 * in the order added, geoms name (tree, link), and one contact system
 * couples them; such models step on the lane-per-world FP64 step in every
 * precision, and under the penalty solver pairs between two trees are
 * rejected at finalize. */
newt_status newt_model_add_tree(newt_model* model, const newt_link_desc* links,
                                int32_t n_links, int32_t self_collide,
                                int32_t* tree_index);

/* World::add_body (newt/src/world.rs:950-954).  A model may hold trees and
 * free bodies together: the bodies' state follows the trees' in qpos/qvel,
 * a contact with a link on one side is solved in the tree solve with the
 * body as a component of its rows (solver.rs:2805-3366), the others in the
 * free-body solve (world.rs:2097-2149).  Such models step on the
 * lane-per-world FP64 step in every precision, under PGS or Newton, with
 * state sensors on links and bodies; penalty contacts and acceleration,
 * force or touch sensors on them are rejected at finalize with
 * NEWT_ERR_UNSUPPORTED. */
newt_status newt_model_add_body(newt_model* model, const newt_body_desc* body,
                                int32_t* body_index);

/* World::meshes (newt/src/world.rs; ConvexMesh, newt/src/geom.rs:305-340):
 * a convex mesh of n_vertices mesh-local points (x, y, z) and n_faces
 * triangles (three vertex indices each, counter-clockwise seen from
 * outside).  ConvexMesh::validate's checks apply (at least 4 vertices and 4
 * faces, indices in range, finite coordinates): NEWT_ERR_INVALID_ARG
 * otherwise.  Geoms refer to it by index (newt_geom_desc::mesh).  Mesh geoms
 * meet planes (mesh_plane, contact.rs:1424-1510), spheres (sphere_mesh,
 * 1720-1776) and meshes (GJK + EPA and the face/edge manifold,
 * contact/ccd.rs:1594-2092); models with meshes step on the lane-per-world
 * FP64 kernels in every precision. */
newt_status newt_model_add_mesh(newt_model* model, const double* vertices, int32_t n_vertices,
                                const int32_t* faces, int32_t n_faces, int32_t* mesh_index);

/* World::add_geom (newt/src/world.rs:974). */
newt_status newt_model_add_geom(newt_model* model, const newt_geom_desc* geom,
                                int32_t* geom_index);

/* Validate (SolverConfig::validate, assert_pairs_supported) and compile the
 * model into device tables.  No further add_* calls are accepted. */
newt_status newt_model_finalize(newt_model* model);

/* Most contacts per world (a configuration cap such as "up to 512 contacts
 * per world"; newt itself has none): a world whose step finds
 * more is frozen with NEWT_ERR_CAPACITY at its start-of-step state while the
 * other worlds go on.  0 (the default) = the largest count the model's
 * candidate pairs can produce.  Before newt_model_finalize. */
newt_status newt_model_set_contact_capacity(newt_model* model, int32_t max_contacts);

/* A model from newt's JSON scene format (model::load_str,
 * newt/src/model.rs:179-195; docs/model-format.md): strict like newt's
 * loader (unknown fields rejected; newt_last_error names the JSON path of
 * the offending value), numbers rounded to f32 as newt parses them.  What
 * this build does not step (trees with bodies, actuators,
 * equalities, tendons, meshes, heightfields, sensors, explicit pair lists,
 * ball joints) returns NEWT_ERR_INVALID_CONFIG.  finalize = 1 also runs
 * newt_model_finalize; 0 returns the model open for further add_* calls. */
newt_status newt_model_from_json(const char* json, int32_t finalize, newt_model** out);

/* A model's descriptions, as built (before or after finalize): counts,
 * options, one tree's links (up to `cap`, with its link count and
 * self-collision flag), one geom, one body. */
newt_status newt_model_counts(const newt_model* model, int32_t* n_trees, int32_t* n_bodies, int32_t* n_geoms);
newt_status newt_model_get_options(const newt_model* model, newt_options* out);
newt_status newt_model_get_tree(const newt_model* model, int32_t tree, newt_link_desc* links, int32_t cap,
                                int32_t* n_links, int32_t* self_collide);
newt_status newt_model_get_geom(const newt_model* model, int32_t index, newt_geom_desc* out);
newt_status newt_model_get_body(const newt_model* model, int32_t index, newt_body_desc* out);

/* Sizes of one world: nq, nv, link count, geom count.  Free body i holds
 * qpos[7i..7i+7) = (position, orientation x y z w) and qvel[6i..6i+6) =
 * (linear velocity in world, angular velocity in body frame), after the
 * tree's coordinates. */
newt_status newt_model_dims(const newt_model* model, int32_t* nq, int32_t* nv,
                            int32_t* n_links, int32_t* n_geoms);

void newt_model_destroy(newt_model* model);

/* ---- data: the mutable half of World, for n_worlds worlds ------------ */

/* Allocates device state for n_worlds worlds.  Each world starts at the
 * default state of Tree::push_link (newt/src/tree.rs:275-330). */
newt_status newt_data_create(const newt_model* model, int32_t n_worlds,
                             int32_t precision, newt_data** out);

void newt_data_destroy(newt_data* data);

/* Tree::qfrc_applied (tree.rs:181-237), [n_worlds*nv]: the user's
 * generalized force, held across steps.  A step adds its constraint force
 * to it and takes it back off (world.rs:1899-1946, 2020-2049), so it may
 * change by rounding as newt's does. */
newt_status newt_data_set_qfrc_applied(newt_data* data, const double* qfrc);
newt_status newt_data_get_qfrc_applied(const newt_data* data, double* qfrc);

/* newt_data_get_faults value of a parked world slot (see
 * newt_data_park_worlds); not a newt_status. */
#define NEWT_WORLD_PARKED (-1)

/* Host copies of the generalized state: qpos [n_worlds*nq], qvel
 * [n_worlds*nv].  The getters (and set_state) are ordered after the work
 * last enqueued for this data and wait only for that stream.  set_state also rebuilds each world's broadphase fat
 * bounds from the new state, as World::add_geom and the public state
 * setters do (newt/src/world.rs:974-979, 1119-1136), and clears faults. */
newt_status newt_data_set_state(newt_data* data, const double* qpos,
                                const double* qvel);
newt_status newt_data_get_state(const newt_data* data, double* qpos,
                                double* qvel);

/* set_state enqueued on `stream` (a hipStream_t; NULL = the default
 * stream): the host arrays are copied into a pinned staging buffer of the
 * data before the call returns and are free to reuse; the device copy and
 * the fault reset run in stream order, so steps enqueued later on `stream`
 * see the new state.  For serving loops that must not wait on other
 * streams.  Same semantics as newt_data_set_state (world.rs:1119-1136). */
newt_status newt_data_set_state_async(newt_data* data, const double* qpos,
                                      const double* qvel, void* stream);

/* World slots, for serving loops that batch many requests into one
 * newt_data.  set_worlds_async starts `count` new worlds in slots
 * [first, first+count) from qpos [count*nq] / qvel [count*nv], stream-ordered
 * like newt_data_set_state_async: zero applied force, fresh broadphase
 * proxies (World::new + add_geom + set state, world.rs:326-361,974-979),
 * faults cleared.  park_worlds marks slots NEWT_WORLD_PARKED: newt_step
 * leaves them untouched until set_worlds_async reuses them.  get_worlds
 * copies a slot range's state out after the work enqueued for this data.
 * No counterpart in newt, which has one World per object. */
newt_status newt_data_set_worlds_async(newt_data* data, int32_t first, int32_t count,
                                       const double* qpos, const double* qvel, void* stream);
newt_status newt_data_park_worlds(newt_data* data, int32_t first, int32_t count, void* stream);
newt_status newt_data_get_worlds(const newt_data* data, int32_t first, int32_t count,
                                 double* qpos, double* qvel);

/* Worlds with world_mask[w] != 0 (every world when world_mask is NULL) go
 * back to the model's default state, as a fresh World (World::new +
 * add_tree/add_body/add_geom, newt/src/world.rs:326-361, 950-979): default
 * qpos/qvel, zero applied force, proxies at the default state, faults
 * cleared.  Stream-ordered like newt_data_set_state_async. */
newt_status newt_data_reset(newt_data* data, const uint8_t* world_mask, void* stream);

/* World::reset_to_keyframe (newt/src/world.rs:492-519) on the masked worlds:
 * qpos [nq] and qvel [nv] (one keyframe, every masked world) replace the
 * state, the applied force is kept, and the proxies are refit from the new
 * state at the next step (refresh_broadphase_for_query); faults cleared. */
newt_status newt_data_reset_to_keyframe(newt_data* data, const uint8_t* world_mask, const double* qpos,
                                        const double* qvel, void* stream);

/* A device copy of a data's state (qpos, qvel, qacc, applied force, proxy
 * bounds, faults, contact and pair counts): the batched form of World's
 * Clone (newt/src/world.rs:107).  newt_data_snapshot waits for the data's
 * last stream; restore is enqueued on `stream` and needs a snapshot of a
 * data of the same model and world count.  The last step's contact and
 * pair lists are not part of it. */
typedef struct newt_snapshot newt_snapshot;
newt_status newt_data_snapshot(const newt_data* data, newt_snapshot** out);
newt_status newt_data_restore(newt_data* data, const newt_snapshot* snapshot, void* stream);
void newt_snapshot_destroy(newt_snapshot* snapshot);

/* Device memory allocated by this data, in bytes (admission control). */
newt_status newt_data_device_bytes(const newt_data* data, uint64_t* bytes);

/* Generalized acceleration of the last step's integration: the qddot that
 * euler_step_with_workspace adds to qdot (newt/src/tree.rs:1407-1417).
 * [n_worlds*nv]. */
newt_status newt_data_get_qacc(const newt_data* data, double* qacc);

/* Per-world fault flags [n_worlds] (0 = healthy; else a newt_status). */
newt_status newt_data_get_faults(const newt_data* data, int32_t* faults);

/* Contacts consumed by the last step's solver, in narrow-phase order
 * (newt/src/world.rs:2476-2528): for world w, counts[w] contacts; entry
 * [w*cap + k] holds (geom_a, geom_b, slot) in `ids` (3 int32 each) and
 * (position xyz, normal xyz, penetration) in `geo` (7 doubles each).
 * `cap` is the per-world capacity reported by newt_data_contact_capacity.
 * Under the penalty solver (trees or free bodies), RK4 reports the contacts
 * of the step's last stage (collect_contacts, world.rs:2434-2474, at that
 * stage's state); Euler and ImplicitFast report the start-of-step contacts.
 * A world frozen by a fault keeps its last completed state. */
newt_status newt_data_contact_capacity(const newt_data* data, int32_t* cap);
newt_status newt_data_get_contacts(const newt_data* data, int32_t* counts,
                                   int32_t* ids, double* geo);

/* Broadphase candidate pairs of the last step (World::active_pairs,
 * newt/src/world.rs:1173-1226), sorted (min, max): counts[w] pairs, entry
 * [w*pair_cap + k] = (a, b). */
newt_status newt_data_pair_capacity(const newt_data* data, int32_t* cap);
newt_status newt_data_get_pairs(const newt_data* data, int32_t* counts,
                                int32_t* pairs);

/* ---- sensors (newt/src/sensor.rs) ---------------------------------------- */

/* SensorKind (newt/src/sensor.rs:153-196), numbered in newt's declaration
 * order; the kinds this build evaluates. */
typedef enum newt_sensor_kind {
    NEWT_SENSOR_JOINT_POS = 0,      /* hinge angle / slide displacement, 1 */
    NEWT_SENSOR_JOINT_VEL = 1,      /* hinge / slide rate, 1 */
    NEWT_SENSOR_FRAME_POS = 4,      /* site position in world, 3 */
    NEWT_SENSOR_FRAME_QUAT = 5,     /* site orientation in world (x, y, z, w), 4 */
    NEWT_SENSOR_GYRO = 6,           /* parent angular velocity in the site frame, 3 */
    NEWT_SENSOR_ACCELEROMETER = 7,  /* specific force at the site, site frame, 3 */
    NEWT_SENSOR_TOUCH = 8,          /* sum of contact normal-force magnitudes on a geom, 1 */
    NEWT_SENSOR_FORCE = 9,          /* parent-joint force on a link, child frame, 3 */
    NEWT_SENSOR_TORQUE = 10,        /* parent-joint torque at the joint anchor, child frame, 3 */
    NEWT_SENSOR_VELOCIMETER = 13,   /* site linear velocity in the site frame, 3 */
    NEWT_SENSOR_MAGNETOMETER = 14,  /* World::magnetic_field in the site frame, 3 */
    NEWT_SENSOR_SUBTREE_COM = 16,   /* centre of mass of a link's subtree, world, 3 */
    NEWT_SENSOR_FRAME_LINVEL = 17,  /* site linear velocity in world, 3 */
    NEWT_SENSOR_FRAME_ANGVEL = 18   /* parent angular velocity in world, 3 */
} newt_sensor_kind;

/* One Sensor (newt/src/sensor.rs:226-229) with its SiteFrame (sensor.rs:118-122).
 * Joint, force, torque and subtree sensors name a link (tree, link); site
 * sensors name the link or the free body the site rides on (the other
 * index -1), with the site's offset in that body's COM frame and the
 * orientation mapping site vectors into it (x, y, z, w); touch names a geom. */
typedef struct newt_sensor_desc {
    int32_t kind;     /* newt_sensor_kind */
    int32_t tree;     /* the link's tree (SensorKind's `tree`) */
    int32_t link;     /* link of that tree, or -1 */
    int32_t body;     /* free body, or -1 */
    int32_t geom;     /* touch: the geom, else -1 */
    double offset[3];
    double quat[4];
} newt_sensor_desc;

/* World::add_sensor (newt/src/world.rs:441; SensorBank::push, sensor.rs:395-401):
 * the sensor's values follow the earlier sensors' in sensordata, in
 * declaration order.  Accelerometer, force, torque and touch need the
 * post-step accelerations and contact forces; this build evaluates them on
 * tree models under the PGS solver (the step's start-of-step contacts
 * re-solved at the post-step state, World::build_sensor_inputs,
 * world.rs:1578-1696) or on tree models without contact pairs; elsewhere
 * finalize returns NEWT_ERR_UNSUPPORTED.  Before newt_model_finalize. */
newt_status newt_model_add_sensor(newt_model* model, const newt_sensor_desc* sensor, int32_t* sensor_index);
/* World::magnetic_field (newt/src/world.rs:119), read by magnetometers.
 * Before finalize. */
newt_status newt_model_set_magnetic_field(newt_model* model, const double field[3]);
/* Values per world (SensorBank::data length). */
newt_status newt_model_sensor_dim(const newt_model* model, int32_t* dim);
/* World::sensordata after the last newt_step (sensor::evaluate on the
 * post-step state, sensor.rs:446-702), [n_worlds * dim]; zeros before the
 * first step.  newt evaluates after every step and keeps the last values;
 * here they are evaluated once after each newt_step call, which gives the
 * same values.  Worlds frozen by a fault keep their previous values. */
newt_status newt_data_get_sensordata(const newt_data* data, double* sensordata);

/* ---- step --------------------------------------------------------------- */

/* n_steps × World::step (newt/src/world.rs:1459) on every world of `data`,
 * enqueued on `stream` (a hipStream_t; NULL = the default stream).  Returns
 * once the work is enqueued; faults are read with newt_data_get_faults. */
newt_status newt_step(const newt_model* model, newt_data* data,
                      int32_t n_steps, void* stream);

/* Blocks until every step enqueued on `data` has finished (waits on the
 * stream of its last step). */
newt_status newt_data_synchronize(newt_data* data);

/* ---- uneven batches: several datas in one launch ------------------------- */

/* A batch of newt_datas of different models (e.g. trees of different dof
 * counts) stepped together.  No counterpart in newt, which steps one World
 * per call (newt/src/world.rs:1459-1537); each world's step is the same as
 * newt_step's, bit for bit.  The datas are grouped by kernel instantiation
 * (precision, register width for nv <= 24 / 32 / 64, extended step); each
 * group is one launch of the persistent step over a table of jobs, one job
 * (a run of workgroups) per data.  The datas must run the persistent step:
 * tree models, NEWT_PRECISION_FP64 or _MIXED, no penalty joint limits. */
typedef struct newt_batch newt_batch;

typedef struct newt_batch_options {
    int32_t order;        /* 1: datas with the most dofs first (they bound the
                             launch); 0: as given */
    int32_t wpb_cap[3];   /* most worlds per workgroup for register widths 24,
                             32, 64 (0: as many as the LDS holds) */
    int32_t use_graph;    /* 1: newt_batch_step replays a HIP graph of the
                             group launches, captured once per step count */
    int32_t branches;     /* 1: the groups run on parallel graph branches */
} newt_batch_options;

/* `options` NULL: order 1, no caps, use_graph 1, branches 1.  The datas stay
 * owned by the caller and must outlive the batch. */
newt_status newt_batch_create(newt_data* const* datas, int32_t n_datas, const newt_batch_options* options,
                              newt_batch** out);
/* n_steps World::steps on every world of every data of the batch, enqueued on
 * `stream` (a hipStream_t; NULL = the default stream).  The datas' getters
 * are ordered after it. */
newt_status newt_batch_step(newt_batch* batch, int32_t n_steps, void* stream);
/* Kernel launches per newt_batch_step (the number of groups). */
newt_status newt_batch_launches(const newt_batch* batch, int32_t* n_launches);
void newt_batch_destroy(newt_batch* batch);

/* ---- step without allocation --------------------------------------------- */

/* Device state of n_worlds worlds owned by the caller, lane-interleaved:
 * element k of world w at [k * n_worlds + w].  Same meaning as a newt_data's
 * state: qpos [nq], qvel/qacc/qfrc [nv] (qfrc = Tree::qfrc_applied), fat
 * [6 * n_geoms] broadphase proxy bounds (min xyz, max xyz per geom), and per
 * world the fault code, the last step's contact and pair counts. */
typedef struct newt_device_state {
    double* qpos;
    double* qvel;
    double* qacc;
    double* qfrc;
    double* fat;
    int32_t* fault;
    int32_t* ncontact;
    int32_t* npair;
} newt_device_state;

/* newt_step on caller-owned device state, allocating nothing: no hipMalloc,
 * hipFree or other allocating call; its scratch (the last step's contacts
 * and pairs, and the Jacobian rows of contact sets over 64 rows) is a
 * static device array of 160 MiB per GPU, held until the process ends.
 * No counterpart in the reference (newt steps one World in host memory);
 * it is for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation.
 *   reset_bounds = 1 rebuilds every proxy bound from qpos/qvel before the
 *   first step, as newt_data_set_state does; 0 continues from `fat`.
 * Tree models, NEWT_PRECISION_FP64 or _MIXED; other models return
 * NEWT_ERR_UNSUPPORTED.  A batch over the scratch size returns
 * NEWT_ERR_CAPACITY (newt_noalloc_max_worlds gives the largest).  The
 * scratch keeps the last call's contents; calls that overlap on one GPU
 * race. */
newt_status newt_step_noalloc(const newt_model* model, const newt_device_state* state, int32_t n_worlds,
                              int32_t precision, int32_t reset_bounds, int32_t n_steps, void* stream);
newt_status newt_noalloc_max_worlds(const newt_model* model, int32_t precision, int32_t* max_worlds);

#ifdef __cplusplus
}
#endif

#endif /* NEWT_GPU_H */
