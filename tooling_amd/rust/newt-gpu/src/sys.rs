//! Raw bindings of include/newt_gpu.h.  Keep the #[repr(C)] layouts in sync
//! with the header (tests/test_abi.py checks the header's layouts from C).
#![allow(dead_code)]
use std::os::raw::{c_char, c_void};

pub type NewtStatus = i32;
pub const NEWT_OK: NewtStatus = 0;
pub const NEWT_ERR_INVALID_CONFIG: NewtStatus = 1;
pub const NEWT_ERR_UNSUPPORTED_PAIR: NewtStatus = 2;
pub const NEWT_ERR_NOT_PD: NewtStatus = 3;
pub const NEWT_ERR_NEWTON_FAIL: NewtStatus = 4;
pub const NEWT_ERR_OOM: NewtStatus = 5;
pub const NEWT_ERR_DEVICE: NewtStatus = 6;
pub const NEWT_ERR_INVALID_ARG: NewtStatus = 7;
pub const NEWT_ERR_CAPACITY: NewtStatus = 8;
pub const NEWT_ERR_UNSUPPORTED: NewtStatus = 9;
pub const NEWT_WORLD_PARKED: i32 = -1;

pub const NEWT_JOINT_FREE: i32 = 0;
pub const NEWT_JOINT_FIXED: i32 = 1;
pub const NEWT_JOINT_HINGE: i32 = 2;
pub const NEWT_JOINT_SLIDE: i32 = 3;
pub const NEWT_JOINT_BALL: i32 = 4;
pub const NEWT_SHAPE_PLANE: i32 = 0;
pub const NEWT_SHAPE_SPHERE: i32 = 1;
pub const NEWT_SHAPE_BOX: i32 = 2;
pub const NEWT_SHAPE_CAPSULE: i32 = 3;
pub const NEWT_SHAPE_MESH: i32 = 4;
pub const NEWT_INTEGRATOR_EULER: i32 = 0;
pub const NEWT_INTEGRATOR_IMPLICITFAST: i32 = 1;
pub const NEWT_INTEGRATOR_RK4: i32 = 2;
pub const NEWT_SOLVER_PENALTY: i32 = 0;
pub const NEWT_SOLVER_PGS: i32 = 1;
pub const NEWT_SOLVER_NEWTON: i32 = 2;
pub const NEWT_CONE_PYRAMIDAL: i32 = 0;
pub const NEWT_CONE_ELLIPTIC: i32 = 1;
pub const NEWT_PRECISION_FP64: i32 = 0;
pub const NEWT_PRECISION_MIXED: i32 = 1;
pub const NEWT_PRECISION_FP64_EXACT: i32 = 2;

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct NewtOptions {
    pub dt: f64,
    pub gravity: [f64; 3],
    pub penetration_slop: f64,
    pub integrator: i32,
    pub solver_mode: i32,
    pub iterations: i32,
    pub cone: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct NewtLinkDesc {
    pub parent: i32,
    pub joint: i32,
    pub axis: [f64; 3],
    pub damping: f64,
    pub armature: f64,
    pub free_damping: f64,
    pub has_range: i32,
    pub pad0: i32,
    pub range: [f64; 2],
    pub parent_pos: [f64; 3],
    pub parent_quat: [f64; 4],
    pub child_pos: [f64; 3],
    pub child_quat: [f64; 4],
    pub mass: f64,
    pub inertia: [f64; 9], // column-major, like Mat3
    pub limit_custom: i32,
    pub limit_solimp_power: i32,
    pub limit_solref: [f64; 2],
    pub limit_solimp: [f64; 4],
    pub limit_stiffness: f64,
    pub limit_damping: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct NewtBodyDesc {
    pub mass: f64,
    pub inertia: [f64; 9],
    pub position: [f64; 3],
    pub orientation: [f64; 4],
    pub linear_velocity: [f64; 3],
    pub angular_velocity_body: [f64; 3],
    pub has_max_linear_velocity: i32,
    pub has_max_angular_velocity: i32,
    pub max_linear_velocity: f64,
    pub max_angular_velocity: f64,
    pub gravity_scale: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct NewtSensorDesc {
    pub kind: i32,
    pub tree: i32,
    pub link: i32,
    pub body: i32,
    pub geom: i32,
    pub offset: [f64; 3],
    pub quat: [f64; 4],
}

// newt_sensor_kind: newt's SensorKind declaration order.
pub const NEWT_SENSOR_JOINT_POS: i32 = 0;
pub const NEWT_SENSOR_JOINT_VEL: i32 = 1;
pub const NEWT_SENSOR_FRAME_POS: i32 = 4;
pub const NEWT_SENSOR_FRAME_QUAT: i32 = 5;
pub const NEWT_SENSOR_GYRO: i32 = 6;
pub const NEWT_SENSOR_ACCELEROMETER: i32 = 7;
pub const NEWT_SENSOR_TOUCH: i32 = 8;
pub const NEWT_SENSOR_FORCE: i32 = 9;
pub const NEWT_SENSOR_TORQUE: i32 = 10;
pub const NEWT_SENSOR_VELOCIMETER: i32 = 13;
pub const NEWT_SENSOR_MAGNETOMETER: i32 = 14;
pub const NEWT_SENSOR_SUBTREE_COM: i32 = 16;
pub const NEWT_SENSOR_FRAME_LINVEL: i32 = 17;
pub const NEWT_SENSOR_FRAME_ANGVEL: i32 = 18;

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct NewtGeomDesc {
    pub shape: i32,
    pub tree: i32,
    pub link: i32,
    pub condim: i32,
    pub size: [f64; 3],
    pub pos: [f64; 3],
    pub quat: [f64; 4],
    pub friction: f64,
    pub margin: f64,
    pub gap: f64,
    pub solref: [f64; 2],
    pub solimp: [f64; 4],
    pub solimp_power: i32,
    pub group: u32,
    pub mask: u32,
    pub body: i32,
    pub mesh: i32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct NewtDeviceState {
    pub qpos: *mut f64,
    pub qvel: *mut f64,
    pub qacc: *mut f64,
    pub qfrc: *mut f64,
    pub fat: *mut f64,
    pub fault: *mut i32,
    pub ncontact: *mut i32,
    pub npair: *mut i32,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct NewtBatchOptions {
    pub order: i32,
    pub wpb_cap: [i32; 3],
    pub use_graph: i32,
    pub branches: i32,
}

pub enum NewtModel {}
pub enum NewtData {}
pub enum NewtBatch {}
pub enum NewtSnapshot {}

extern "C" {
    pub fn newt_last_error() -> *const c_char;

    pub fn newt_model_create(o: *const NewtOptions, out: *mut *mut NewtModel) -> NewtStatus;
    pub fn newt_model_add_tree(m: *mut NewtModel, links: *const NewtLinkDesc, n: i32, self_collide: i32,
                               idx: *mut i32) -> NewtStatus;
    pub fn newt_model_add_body(m: *mut NewtModel, b: *const NewtBodyDesc, idx: *mut i32) -> NewtStatus;
    pub fn newt_model_add_geom(m: *mut NewtModel, g: *const NewtGeomDesc, idx: *mut i32) -> NewtStatus;
    pub fn newt_model_add_mesh(m: *mut NewtModel, vertices: *const f64, n_vertices: i32, faces: *const i32,
                               n_faces: i32, idx: *mut i32) -> NewtStatus;
    pub fn newt_model_set_contact_capacity(m: *mut NewtModel, max_contacts: i32) -> NewtStatus;
    pub fn newt_model_finalize(m: *mut NewtModel) -> NewtStatus;
    pub fn newt_model_from_json(json: *const c_char, finalize: i32, out: *mut *mut NewtModel) -> NewtStatus;
    pub fn newt_model_counts(m: *const NewtModel, n_trees: *mut i32, n_bodies: *mut i32, n_geoms: *mut i32)
                             -> NewtStatus;
    pub fn newt_model_get_options(m: *const NewtModel, out: *mut NewtOptions) -> NewtStatus;
    pub fn newt_model_get_tree(m: *const NewtModel, tree: i32, links: *mut NewtLinkDesc, cap: i32,
                               n_links: *mut i32, self_collide: *mut i32) -> NewtStatus;
    pub fn newt_model_get_geom(m: *const NewtModel, index: i32, out: *mut NewtGeomDesc) -> NewtStatus;
    pub fn newt_model_get_body(m: *const NewtModel, index: i32, out: *mut NewtBodyDesc) -> NewtStatus;
    pub fn newt_model_dims(m: *const NewtModel, nq: *mut i32, nv: *mut i32, n_links: *mut i32, n_geoms: *mut i32)
                           -> NewtStatus;
    pub fn newt_model_destroy(m: *mut NewtModel);

    pub fn newt_data_create(m: *const NewtModel, n_worlds: i32, precision: i32, out: *mut *mut NewtData)
                            -> NewtStatus;
    pub fn newt_data_destroy(d: *mut NewtData);
    pub fn newt_data_set_qfrc_applied(d: *mut NewtData, qfrc: *const f64) -> NewtStatus;
    pub fn newt_data_get_qfrc_applied(d: *const NewtData, qfrc: *mut f64) -> NewtStatus;
    pub fn newt_data_set_state(d: *mut NewtData, qpos: *const f64, qvel: *const f64) -> NewtStatus;
    pub fn newt_data_get_state(d: *const NewtData, qpos: *mut f64, qvel: *mut f64) -> NewtStatus;
    pub fn newt_data_set_state_async(d: *mut NewtData, qpos: *const f64, qvel: *const f64, stream: *mut c_void)
                                     -> NewtStatus;
    pub fn newt_data_set_worlds_async(d: *mut NewtData, first: i32, count: i32, qpos: *const f64,
                                      qvel: *const f64, stream: *mut c_void) -> NewtStatus;
    pub fn newt_data_park_worlds(d: *mut NewtData, first: i32, count: i32, stream: *mut c_void) -> NewtStatus;
    pub fn newt_data_get_worlds(d: *const NewtData, first: i32, count: i32, qpos: *mut f64, qvel: *mut f64)
                                -> NewtStatus;
    pub fn newt_data_reset(d: *mut NewtData, world_mask: *const u8, stream: *mut c_void) -> NewtStatus;
    pub fn newt_data_reset_to_keyframe(d: *mut NewtData, world_mask: *const u8, qpos: *const f64,
                                       qvel: *const f64, stream: *mut c_void) -> NewtStatus;
    pub fn newt_data_snapshot(d: *const NewtData, out: *mut *mut NewtSnapshot) -> NewtStatus;
    pub fn newt_data_restore(d: *mut NewtData, s: *const NewtSnapshot, stream: *mut c_void) -> NewtStatus;
    pub fn newt_snapshot_destroy(s: *mut NewtSnapshot);
    pub fn newt_data_device_bytes(d: *const NewtData, bytes: *mut u64) -> NewtStatus;
    pub fn newt_data_get_qacc(d: *const NewtData, qacc: *mut f64) -> NewtStatus;
    pub fn newt_data_get_faults(d: *const NewtData, faults: *mut i32) -> NewtStatus;
    pub fn newt_data_contact_capacity(d: *const NewtData, cap: *mut i32) -> NewtStatus;
    pub fn newt_data_get_contacts(d: *const NewtData, counts: *mut i32, ids: *mut i32, geo: *mut f64)
                                  -> NewtStatus;
    pub fn newt_data_pair_capacity(d: *const NewtData, cap: *mut i32) -> NewtStatus;
    pub fn newt_data_get_pairs(d: *const NewtData, counts: *mut i32, pairs: *mut i32) -> NewtStatus;

    pub fn newt_model_add_sensor(m: *mut NewtModel, s: *const NewtSensorDesc, index: *mut i32) -> NewtStatus;
    pub fn newt_model_set_magnetic_field(m: *mut NewtModel, field: *const f64) -> NewtStatus;
    pub fn newt_model_sensor_dim(m: *const NewtModel, dim: *mut i32) -> NewtStatus;
    pub fn newt_data_get_sensordata(d: *const NewtData, out: *mut f64) -> NewtStatus;
    pub fn newt_step(m: *const NewtModel, d: *mut NewtData, n_steps: i32, stream: *mut c_void) -> NewtStatus;
    pub fn newt_data_synchronize(d: *mut NewtData) -> NewtStatus;

    pub fn newt_batch_create(datas: *const *mut NewtData, n: i32, opt: *const NewtBatchOptions,
                             out: *mut *mut NewtBatch) -> NewtStatus;
    pub fn newt_batch_step(b: *mut NewtBatch, n_steps: i32, stream: *mut c_void) -> NewtStatus;
    pub fn newt_batch_launches(b: *const NewtBatch, n: *mut i32) -> NewtStatus;
    pub fn newt_batch_destroy(b: *mut NewtBatch);

    pub fn newt_step_noalloc(m: *const NewtModel, state: *const NewtDeviceState, n_worlds: i32, precision: i32,
                             reset_bounds: i32, n_steps: i32, stream: *mut c_void) -> NewtStatus;
    pub fn newt_noalloc_max_worlds(m: *const NewtModel, precision: i32, max_worlds: *mut i32) -> NewtStatus;
}
