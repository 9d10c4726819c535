//! newt's Model / Data / step() over the MI355X step (libnewt_gpu.so).
//!
//! newt keeps model and state in one `World` and advances it with
//! `World::step(&mut self)` (newt/src/world.rs:1459-1537).  Here a [`Model`]
//! is the immutable half of a `World` (options, trees, bodies, geoms),
//! compiled once into device tables, and a [`Data`] holds the state of N
//! worlds on the GPU; [`step`] advances every world of a `Data` by n
//! `World::step` calls.  Errors are `Err(NewtError)` where newt panics.
//! A panic-class condition inside one world (non-PD mass matrix, a contact
//! set over the capacity) freezes that world with a fault code; the others
//! go on.
//!
//! This crate is the binding a newt maintainer would add; the repository
//! that ships it builds no Rust, and the tested boundary is the C ABI it
//! wraps (include/newt_gpu.h).
pub mod sys;

use std::ffi::{CStr, CString};
use std::os::raw::c_void;

use newt::geom::{Geom, GeomShape};
use newt::joint::JointKind;
use newt::solver::{ConeKind, SolverMode};
use newt::world::{Integrator, World};
use sys::*;

/// A failed call: the newt_status and newt_last_error()'s text.
#[derive(Debug, Clone)]
pub struct NewtError {
    pub status: NewtStatus,
    pub message: String,
}

impl std::fmt::Display for NewtError {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "newt_gpu status {}: {}", self.status, self.message)
    }
}
impl std::error::Error for NewtError {}

fn check(s: NewtStatus) -> Result<(), NewtError> {
    if s == NEWT_OK {
        return Ok(());
    }
    let message = unsafe { CStr::from_ptr(newt_last_error()) }.to_string_lossy().into_owned();
    Err(NewtError { status: s, message })
}

/// Arithmetic of a [`Data`].
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Precision {
    /// The drop-in: FP64, within 1e-9 of the CPU step.
    Fp64 = NEWT_PRECISION_FP64 as isize,
    /// FP32 storage of the Jacobian rows; qacc within 1e-5 of the FP64 step.
    Mixed = NEWT_PRECISION_MIXED as isize,
    /// FP64 in newt's operation order: bit-identical to the CPU restatement.
    Fp64Exact = NEWT_PRECISION_FP64_EXACT as isize,
}

/// The immutable half of a `World`, compiled for the device.
pub struct Model {
    raw: *mut NewtModel,
    pub nq: usize,
    pub nv: usize,
}

fn v3(v: newt::math::Vec3) -> [f64; 3] {
    [v.x as f64, v.y as f64, v.z as f64]
}
fn q4(q: newt::math::Quat) -> [f64; 4] {
    [q.x as f64, q.y as f64, q.z as f64, q.w as f64]
}
fn m9(m: &newt::math::Mat3) -> [f64; 9] {
    // Mat3 is column-major (math.rs:386).
    let mut out = [0.0; 9];
    for (k, x) in m.to_cols_array().iter().enumerate() {
        out[k] = *x as f64;
    }
    out
}

/// One Link and its joint (tree.rs:79-123, joint.rs:40-165).
fn lower_link(l: &newt::tree::Link) -> NewtLinkDesc {
    let mut d = NewtLinkDesc { parent: l.parent.map(|p| p as i32).unwrap_or(-1), ..Default::default() };
    d.free_damping = l.free_damping as f64;
    d.parent_pos = v3(l.joint_offset_in_parent.0);
    d.parent_quat = q4(l.joint_offset_in_parent.1);
    d.child_pos = v3(l.joint_offset_in_child.0);
    d.child_quat = q4(l.joint_offset_in_child.1);
    d.mass = l.mass as f64;
    d.inertia = m9(&l.inertia_body);
    match &l.joint {
        JointKind::Free => d.joint = NEWT_JOINT_FREE,
        JointKind::Fixed => d.joint = NEWT_JOINT_FIXED,
        JointKind::Ball { damping, armature } => {
            d.joint = NEWT_JOINT_BALL; // rejected at finalize: not built
            d.damping = *damping as f64;
            d.armature = *armature as f64;
        }
        JointKind::Hinge { axis, range, damping, armature, limit }
        | JointKind::Slide { axis, range, damping, armature, limit } => {
            d.joint = if matches!(l.joint, JointKind::Hinge { .. }) { NEWT_JOINT_HINGE } else { NEWT_JOINT_SLIDE };
            d.axis = v3(*axis);
            d.damping = *damping as f64;
            d.armature = *armature as f64;
            if let Some((lo, hi)) = range {
                d.has_range = 1;
                d.range = [*lo as f64, *hi as f64];
            }
            d.limit_custom = 1;
            d.limit_stiffness = limit.stiffness as f64;
            d.limit_damping = limit.damping as f64;
            let sr = limit.solref.unwrap_or(newt::geom::SolRef::DEFAULT);
            d.limit_solref = [sr.timeconst as f64, sr.dampratio as f64];
            let si = limit.solimp.unwrap_or(newt::solver::SolImp::DEFAULT);
            d.limit_solimp = [si.dmin as f64, si.dmax as f64, si.width as f64, si.midpoint as f64];
            d.limit_solimp_power = si.power as i32;
        }
    }
    d
}

/// One Geom (geom.rs:352-419); shapes this build does not step keep their
/// kind out of range and are rejected at finalize.
fn lower_geom(g: &Geom) -> NewtGeomDesc {
    let mut d = NewtGeomDesc { tree: -1, link: -1, body: -1, ..Default::default() };
    match &g.shape {
        GeomShape::Plane => d.shape = NEWT_SHAPE_PLANE,
        GeomShape::Sphere { radius } => {
            d.shape = NEWT_SHAPE_SPHERE;
            d.size[0] = *radius as f64;
        }
        GeomShape::Box { half_extents } => {
            d.shape = NEWT_SHAPE_BOX;
            d.size = v3(*half_extents);
        }
        GeomShape::Capsule { radius, half_height } => {
            d.shape = NEWT_SHAPE_CAPSULE;
            d.size = [*radius as f64, *half_height as f64, 0.0];
        }
        GeomShape::Mesh { mesh_id } => {
            d.shape = NEWT_SHAPE_MESH;
            d.mesh = *mesh_id as i32;
        }
        _ => d.shape = -1,
    }
    if let Some((t, l)) = g.link {
        d.tree = t as i32;
        d.link = l as i32;
    } else if let Some(b) = g.body {
        d.body = b as i32;
    }
    d.condim = g.condim as i32;
    d.pos = v3(g.local_offset);
    d.quat = q4(g.local_orientation);
    d.friction = g.friction as f64;
    d.margin = g.margin as f64;
    d.gap = g.gap as f64;
    d.solref = [g.solref.timeconst as f64, g.solref.dampratio as f64];
    d.solimp = [g.solimp.dmin as f64, g.solimp.dmax as f64, g.solimp.width as f64, g.solimp.midpoint as f64];
    d.solimp_power = g.solimp.power as i32;
    d.group = g.collision_group;
    d.mask = g.collision_mask;
    d
}

/// One free Body (body.rs:22-47); the library inverts the inertia as
/// Body::new does.
fn lower_body(b: &newt::body::Body) -> NewtBodyDesc {
    NewtBodyDesc {
        mass: b.mass as f64,
        inertia: m9(&b.inertia_body),
        position: v3(b.position),
        orientation: q4(b.orientation),
        linear_velocity: v3(b.linear_velocity),
        angular_velocity_body: v3(b.angular_velocity_body),
        has_max_linear_velocity: b.max_linear_velocity.is_some() as i32,
        has_max_angular_velocity: b.max_angular_velocity.is_some() as i32,
        max_linear_velocity: b.max_linear_velocity.unwrap_or(0.0) as f64,
        max_angular_velocity: b.max_angular_velocity.unwrap_or(0.0) as f64,
        gravity_scale: b.gravity_scale as f64,
    }
}

impl Model {
    /// Compile `world`'s model (SolverConfig::validate and
    /// assert_pairs_supported run here, as newt's first step would).
    pub fn from_world(world: &World) -> Result<Self, NewtError> {
        let opts = NewtOptions {
            dt: world.dt as f64,
            gravity: v3(world.gravity),
            penetration_slop: world.penetration_slop as f64,
            integrator: match world.integrator {
                Integrator::Euler => NEWT_INTEGRATOR_EULER,
                Integrator::ImplicitFast => NEWT_INTEGRATOR_IMPLICITFAST,
                Integrator::Rk4 => NEWT_INTEGRATOR_RK4,
            },
            solver_mode: match world.solver.mode {
                SolverMode::Penalty => NEWT_SOLVER_PENALTY,
                SolverMode::Pgs => NEWT_SOLVER_PGS,
                SolverMode::Newton => NEWT_SOLVER_NEWTON,
            },
            iterations: world.solver.iterations as i32,
            cone: match world.solver.cone {
                ConeKind::Pyramidal => NEWT_CONE_PYRAMIDAL,
                _ => NEWT_CONE_ELLIPTIC,
            },
        };
        let mut raw = std::ptr::null_mut();
        check(unsafe { newt_model_create(&opts, &mut raw) })?;
        let m = Model { raw, nq: 0, nv: 0 };
        for (t, tree) in world.trees.iter().enumerate() {
            let links: Vec<NewtLinkDesc> = tree.links.iter().map(lower_link).collect();
            let collide = !world.disabled_self_collision.contains(&t) as i32;
            check(unsafe { newt_model_add_tree(m.raw, links.as_ptr(), links.len() as i32, collide, std::ptr::null_mut()) })?;
        }
        for b in world.bodies.iter() {
            check(unsafe { newt_model_add_body(m.raw, &lower_body(b), std::ptr::null_mut()) })?;
        }
        for mesh in world.meshes.iter() {
            let v: Vec<f64> = mesh.vertices.iter().flat_map(|p| [p.x as f64, p.y as f64, p.z as f64]).collect();
            let f: Vec<i32> = mesh.faces.iter().flat_map(|t| [t[0] as i32, t[1] as i32, t[2] as i32]).collect();
            check(unsafe {
                newt_model_add_mesh(m.raw, v.as_ptr(), mesh.vertices.len() as i32, f.as_ptr(), mesh.faces.len() as i32,
                                    std::ptr::null_mut())
            })?;
        }
        for g in world.geoms.iter() {
            check(unsafe { newt_model_add_geom(m.raw, &lower_geom(g), std::ptr::null_mut()) })?;
        }
        for s in world.sensors.sensors.iter() {
            let d = lower_sensor(&s.kind).ok_or_else(|| NewtError {
                status: NEWT_ERR_UNSUPPORTED,
                message: format!("sensor {:?} is not built", s.name),
            })?;
            check(unsafe { newt_model_add_sensor(m.raw, &d, std::ptr::null_mut()) })?;
        }
        let mf = v3(world.magnetic_field);
        check(unsafe { newt_model_set_magnetic_field(m.raw, mf.as_ptr()) })?;
        m.finalize()
    }

    /// A model from newt's JSON scene text (model::load_str semantics).
    pub fn from_json(text: &str) -> Result<Self, NewtError> {
        let c = CString::new(text).map_err(|e| NewtError { status: NEWT_ERR_INVALID_ARG, message: e.to_string() })?;
        let mut raw = std::ptr::null_mut();
        check(unsafe { newt_model_from_json(c.as_ptr(), 0, &mut raw) })?;
        Model { raw, nq: 0, nv: 0 }.finalize()
    }

    fn finalize(mut self) -> Result<Self, NewtError> {
        check(unsafe { newt_model_finalize(self.raw) })?;
        let (mut nq, mut nv) = (0, 0);
        check(unsafe { newt_model_dims(self.raw, &mut nq, &mut nv, std::ptr::null_mut(), std::ptr::null_mut()) })?;
        self.nq = nq as usize;
        self.nv = nv as usize;
        Ok(self)
    }
}

/// newt's Sensor (sensor.rs:153-196) as the C record; None for
/// kinds this build does not evaluate.
fn lower_sensor(kind: &newt::sensor::SensorKind) -> Option<NewtSensorDesc> {
    use newt::sensor::{SensorAttach, SensorKind as K, SiteFrame};
    let joint = |k: i32, tree: usize, link: usize| NewtSensorDesc {
        kind: k, tree: tree as i32, link: link as i32, body: -1, geom: -1, ..Default::default()
    };
    let site = |k: i32, s: &SiteFrame| {
        let (tree, link, body) = match s.attach {
            SensorAttach::Link(t, l) => (t as i32, l as i32, -1),
            SensorAttach::Body(b) => (0, -1, b as i32),
        };
        NewtSensorDesc { kind: k, tree, link, body, geom: -1, offset: v3(s.local_offset), quat: q4(s.local_orientation) }
    };
    Some(match kind {
        K::JointPos { tree, link } => joint(NEWT_SENSOR_JOINT_POS, *tree, *link),
        K::JointVel { tree, link } => joint(NEWT_SENSOR_JOINT_VEL, *tree, *link),
        K::Force { tree, link } => joint(NEWT_SENSOR_FORCE, *tree, *link),
        K::Torque { tree, link } => joint(NEWT_SENSOR_TORQUE, *tree, *link),
        K::SubtreeCom { tree, link } => joint(NEWT_SENSOR_SUBTREE_COM, *tree, *link),
        K::Touch { geom } => NewtSensorDesc { kind: NEWT_SENSOR_TOUCH, tree: 0, link: -1, body: -1, geom: *geom as i32, ..Default::default() },
        K::FramePos(s) => site(NEWT_SENSOR_FRAME_POS, s),
        K::FrameQuat(s) => site(NEWT_SENSOR_FRAME_QUAT, s),
        K::Gyro(s) => site(NEWT_SENSOR_GYRO, s),
        K::Accelerometer(s) => site(NEWT_SENSOR_ACCELEROMETER, s),
        K::Velocimeter(s) => site(NEWT_SENSOR_VELOCIMETER, s),
        K::Magnetometer(s) => site(NEWT_SENSOR_MAGNETOMETER, s),
        K::FrameLinVel(s) => site(NEWT_SENSOR_FRAME_LINVEL, s),
        K::FrameAngVel(s) => site(NEWT_SENSOR_FRAME_ANGVEL, s),
        _ => return None,
    })
}

impl Drop for Model {
    fn drop(&mut self) {
        unsafe { newt_model_destroy(self.raw) }
    }
}

/// The mutable half of N worlds, on the device.  Single-owner, like
/// `&mut World`; distinct Data may step concurrently on different streams.
pub struct Data {
    raw: *mut NewtData,
    pub n_worlds: usize,
    nq: usize,
    nv: usize,
}

impl Data {
    /// N worlds at the model's default state.
    pub fn new(model: &Model, n_worlds: usize, precision: Precision) -> Result<Self, NewtError> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { newt_data_create(model.raw, n_worlds as i32, precision as i32, &mut raw) })?;
        Ok(Data { raw, n_worlds, nq: model.nq, nv: model.nv })
    }

    /// Every world's state, row-major [world][slot]: tree coordinates, then
    /// per free body (position, orientation) / (velocity, body angular
    /// velocity).  Rebuilds the broadphase proxies and clears faults.
    pub fn set_state(&mut self, qpos: &[f64], qvel: &[f64]) -> Result<(), NewtError> {
        assert_eq!(qpos.len(), self.n_worlds * self.nq);
        assert_eq!(qvel.len(), self.n_worlds * self.nv);
        check(unsafe { newt_data_set_state(self.raw, qpos.as_ptr(), qvel.as_ptr()) })
    }

    /// Copies `world`'s current state into every world of the batch.
    pub fn set_state_from(&mut self, world: &World) -> Result<(), NewtError> {
        let (mut q, mut v) = (Vec::new(), Vec::new());
        for tree in world.trees.iter() {
            q.extend(tree.q.iter().map(|x| *x as f64));
            v.extend(tree.qdot.iter().map(|x| *x as f64));
        }
        for b in world.bodies.iter() {
            q.extend_from_slice(&v3(b.position));
            q.extend_from_slice(&q4(b.orientation));
            v.extend_from_slice(&v3(b.linear_velocity));
            v.extend_from_slice(&v3(b.angular_velocity_body));
        }
        let (q, v) = (q.repeat(self.n_worlds), v.repeat(self.n_worlds));
        self.set_state(&q, &v)
    }

    pub fn state(&self) -> Result<(Vec<f64>, Vec<f64>), NewtError> {
        let (mut q, mut v) = (vec![0.0; self.n_worlds * self.nq], vec![0.0; self.n_worlds * self.nv]);
        check(unsafe { newt_data_get_state(self.raw, q.as_mut_ptr(), v.as_mut_ptr()) })?;
        Ok((q, v))
    }

    /// The last step's generalized acceleration, [world][nv].
    pub fn qacc(&self) -> Result<Vec<f64>, NewtError> {
        let mut a = vec![0.0; self.n_worlds * self.nv];
        check(unsafe { newt_data_get_qacc(self.raw, a.as_mut_ptr()) })?;
        Ok(a)
    }

    /// Tree::qfrc_applied of every world, [world][nv].
    pub fn set_qfrc_applied(&mut self, qfrc: &[f64]) -> Result<(), NewtError> {
        assert_eq!(qfrc.len(), self.n_worlds * self.nv);
        check(unsafe { newt_data_set_qfrc_applied(self.raw, qfrc.as_ptr()) })
    }

    /// World::sensordata of every world after the last step, [world][dim].
    pub fn sensordata(&self, model: &Model) -> Result<Vec<f64>, NewtError> {
        let mut dim = 0;
        check(unsafe { newt_model_sensor_dim(model.raw, &mut dim) })?;
        let mut s = vec![0.0; self.n_worlds * dim as usize];
        check(unsafe { newt_data_get_sensordata(self.raw, s.as_mut_ptr()) })?;
        Ok(s)
    }

    /// Per-world fault codes (0 = healthy, else a newt_status).
    pub fn faults(&self) -> Result<Vec<i32>, NewtError> {
        let mut f = vec![0; self.n_worlds];
        check(unsafe { newt_data_get_faults(self.raw, f.as_mut_ptr()) })?;
        Ok(f)
    }

    /// Masked worlds (all with None) back to the model's default state.
    pub fn reset(&mut self, mask: Option<&[bool]>) -> Result<(), NewtError> {
        let m: Option<Vec<u8>> = mask.map(|m| m.iter().map(|x| *x as u8).collect());
        let p = m.as_ref().map(|m| m.as_ptr()).unwrap_or(std::ptr::null());
        check(unsafe { newt_data_reset(self.raw, p, std::ptr::null_mut()) })?;
        check(unsafe { newt_data_synchronize(self.raw) })
    }

    /// World::reset_to_keyframe on the masked worlds.
    pub fn reset_to_keyframe(&mut self, mask: Option<&[bool]>, qpos: &[f64], qvel: &[f64]) -> Result<(), NewtError> {
        assert_eq!(qpos.len(), self.nq);
        assert_eq!(qvel.len(), self.nv);
        let m: Option<Vec<u8>> = mask.map(|m| m.iter().map(|x| *x as u8).collect());
        let p = m.as_ref().map(|m| m.as_ptr()).unwrap_or(std::ptr::null());
        check(unsafe { newt_data_reset_to_keyframe(self.raw, p, qpos.as_ptr(), qvel.as_ptr(), std::ptr::null_mut()) })?;
        check(unsafe { newt_data_synchronize(self.raw) })
    }

    /// A device copy of the state (the batched World::clone).
    pub fn snapshot(&self) -> Result<Snapshot, NewtError> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { newt_data_snapshot(self.raw, &mut raw) })?;
        Ok(Snapshot { raw })
    }

    pub fn restore(&mut self, s: &Snapshot) -> Result<(), NewtError> {
        check(unsafe { newt_data_restore(self.raw, s.raw, std::ptr::null_mut()) })?;
        check(unsafe { newt_data_synchronize(self.raw) })
    }
}

impl Drop for Data {
    fn drop(&mut self) {
        unsafe { newt_data_destroy(self.raw) }
    }
}

pub struct Snapshot {
    raw: *mut NewtSnapshot,
}

impl Drop for Snapshot {
    fn drop(&mut self) {
        unsafe { newt_snapshot_destroy(self.raw) }
    }
}

/// n_steps × World::step on every world of `data`, enqueued on `stream`
/// (a hipStream_t; null = the default stream).
pub fn step(model: &Model, data: &mut Data, n_steps: u32, stream: *mut c_void) -> Result<(), NewtError> {
    check(unsafe { newt_step(model.raw, data.raw, n_steps as i32, stream) })
}

/// Several Data of different models stepped together: one launch per kernel
/// instantiation, replayed from a HIP graph.
pub struct Batch<'a> {
    raw: *mut NewtBatch,
    _datas: std::marker::PhantomData<&'a mut Data>,
}

impl<'a> Batch<'a> {
    pub fn new(datas: &'a mut [&'a mut Data], options: Option<NewtBatchOptions>) -> Result<Self, NewtError> {
        let raws: Vec<*mut NewtData> = datas.iter().map(|d| d.raw).collect();
        let mut raw = std::ptr::null_mut();
        let opt = options.as_ref().map(|o| o as *const NewtBatchOptions).unwrap_or(std::ptr::null());
        check(unsafe { newt_batch_create(raws.as_ptr(), raws.len() as i32, opt, &mut raw) })?;
        Ok(Batch { raw, _datas: std::marker::PhantomData })
    }

    pub fn step(&mut self, n_steps: u32, stream: *mut c_void) -> Result<(), NewtError> {
        check(unsafe { newt_batch_step(self.raw, n_steps as i32, stream) })
    }
}

impl Drop for Batch<'_> {
    fn drop(&mut self) {
        unsafe { newt_batch_destroy(self.raw) }
    }
}
