"""Scene builders for the benchmark workloads, in newt's model vocabulary.

A :class:`Scene` holds the World options, the tree's links, the geoms and a
seeded initial-state sampler.  The same records feed the HIP library and the
CPU oracle, so both step identical doubles.

Inertia helpers restate newt/src/geom.rs:942-1000 (solid box, solid capsule).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List

import numpy as np

from . import abi

IDENTITY_Q = (0.0, 0.0, 0.0, 1.0)


def solid_box_inertia(mass, he):
    """Column-major diag inertia of a solid box (geom.rs:942-951)."""
    hx2, hy2, hz2 = he[0] * he[0], he[1] * he[1], he[2] * he[2]
    return _diag((mass / 3.0) * (hy2 + hz2), (mass / 3.0) * (hx2 + hz2), (mass / 3.0) * (hx2 + hy2))


def solid_capsule_inertia(mass, r, h):
    """Column-major diag inertia of a solid capsule along local Z (geom.rs:966-981)."""
    vol_cyl = r * r * (2.0 * h)
    vol_caps = (4.0 / 3.0) * r * r * r
    total = vol_cyl + vol_caps
    m_c = mass * (vol_cyl / total)
    m_s = mass - m_c
    i_zz = 0.5 * m_c * r * r + (2.0 / 5.0) * m_s * r * r
    offset = h + 3.0 / 8.0 * r
    hemi = (83.0 / 320.0) * m_s * r * r
    i_xx = (1.0 / 12.0) * m_c * (3.0 * r * r + 4.0 * h * h) + hemi + m_s * offset * offset
    return i_xx, i_zz


def _diag(x, y, z):
    return (x, 0.0, 0.0, 0.0, y, 0.0, 0.0, 0.0, z)


def quat_axis_angle(axis, angle):
    n = math.sqrt(sum(a * a for a in axis))
    s, c = math.sin(angle * 0.5), math.cos(angle * 0.5)
    return (axis[0] / n * s, axis[1] / n * s, axis[2] / n * s, c)


def link(parent, joint, *, axis=(0.0, 0.0, 0.0), parent_pos=(0.0, 0.0, 0.0), parent_quat=IDENTITY_Q,
         child_pos=(0.0, 0.0, 0.0), mass=1.0, inertia=_diag(1.0, 1.0, 1.0), damping=0.0, armature=0.0,
         free_damping=0.0, range=None, limit_solref=None, limit_solimp=None, limit_stiffness=None,
         limit_damping=None):
    """`range` (lo, hi) adds PGS limit rows under PGS and a penalty spring
    under the penalty solver; `limit_solref` (timeconst, dampratio) and
    `limit_solimp` (dmin, dmax, width, midpoint, power) override
    SolRef::DEFAULT / SolImp::DEFAULT for the rows, `limit_stiffness` /
    `limit_damping` the spring's JointLimit::DEFAULT 1000 / 60
    (joint.rs:124-165)."""
    return dict(parent=parent, joint=joint, axis=axis, parent_pos=parent_pos, parent_quat=parent_quat,
                child_pos=child_pos, child_quat=IDENTITY_Q, mass=mass, inertia=inertia, damping=damping,
                armature=armature, free_damping=free_damping, range=range, limit_solref=limit_solref,
                limit_solimp=limit_solimp, limit_stiffness=limit_stiffness, limit_damping=limit_damping)


def geom(shape, *, tree=-1, link=-1, body=-1, size=(0.0, 0.0, 0.0), pos=(0.0, 0.0, 0.0), quat=IDENTITY_Q,
         friction=1.0, margin=0.0, gap=0.0, solref=(0.02, 1.0), solimp=(0.9, 0.95, 0.001, 0.5), power=2,
         condim=3, group=0xFFFFFFFF, mask=0xFFFFFFFF, mesh=-1):
    """Geom defaults follow Geom::static_plane / *_on_link / r#box / mesh
    (geom.rs:465-790); `body` attaches it to a free body; `group`/`mask` are
    collision_group / collision_mask (broadphase.rs:16-18); `mesh` indexes
    Scene.meshes for SHAPE_MESH."""
    return dict(shape=shape, tree=tree, link=link, body=body, size=size, pos=pos, quat=quat, friction=friction,
                margin=margin, gap=gap, solref=solref, solimp=solimp, power=power, condim=condim, group=group,
                mask=mask, mesh=mesh)


def body(mass, inertia, position, orientation=IDENTITY_Q, linear_velocity=(0.0, 0.0, 0.0),
         angular_velocity_body=(0.0, 0.0, 0.0), max_linear_velocity=None, max_angular_velocity=None,
         gravity_scale=1.0):
    """A free Body (body.rs:22-70); `inertia` is a column-major 3x3 about
    the COM in body axes."""
    return dict(mass=mass, inertia=tuple(inertia), position=tuple(position), orientation=tuple(orientation),
                linear_velocity=tuple(linear_velocity), angular_velocity_body=tuple(angular_velocity_body),
                max_linear_velocity=max_linear_velocity, max_angular_velocity=max_angular_velocity,
                gravity_scale=gravity_scale)




@dataclass
class Scene:
    name: str
    options: dict
    links: List[dict]
    geoms: List[dict]
    self_collide: int = 0
    sampler: Callable[[int, int], tuple] = None
    description: str = ""
    extra: dict = field(default_factory=dict)
    bodies: List[dict] = field(default_factory=list)
    # Sensors (World::add_sensor): dicts made by `sensor(...)`, in
    # declaration order; `magnetic_field` is World::magnetic_field.
    sensors: List[dict] = field(default_factory=list)
    magnetic_field: tuple = (0.0, 0.0, 0.0)
    # Further trees of the world (World::trees[1..]): (links, self_collide)
    # each; geoms name their tree by index.
    more_trees: List[tuple] = field(default_factory=list)
    # Convex meshes (World::meshes): (vertices [n, 3], faces [m, 3]) each;
    # mesh geoms name theirs by index.
    meshes: List[tuple] = field(default_factory=list)

    @property
    def all_links(self):
        return list(self.links) + [l for t, _ in self.more_trees for l in t]

    @property
    def nq(self):
        return sum(7 if l["joint"] == abi.JOINT_FREE else 0 if l["joint"] == abi.JOINT_FIXED
                   else 4 if l["joint"] == abi.JOINT_BALL else 1 for l in self.all_links) + 7 * len(self.bodies)

    @property
    def nv(self):
        return sum(6 if l["joint"] == abi.JOINT_FREE else 0 if l["joint"] == abi.JOINT_FIXED
                   else 3 if l["joint"] == abi.JOINT_BALL else 1 for l in self.all_links) + 6 * len(self.bodies)

    def tree_arrays(self):
        """(LinkDesc array, count, self_collide) per tree, in World order."""
        out = []
        if self.links:
            out.append((self.link_array(), len(self.links), self.self_collide))
        for links, collide in self.more_trees:
            out.append((self.link_array(links), len(links), collide))
        return out

    def initial_state(self, n_worlds: int, seed: int):
        return self.sampler(n_worlds, seed)

    # -- ctypes records --------------------------------------------------
    def options_struct(self) -> abi.Options:
        o = self.options
        return abi.Options(dt=o["dt"], gravity=abi.D3(*o["gravity"]), penetration_slop=o.get("slop", 0.0),
                           integrator=o["integrator"], solver_mode=o["solver"], iterations=o["iterations"],
                           cone=o.get("cone", abi.CONE_PYRAMIDAL))

    def link_array(self, links=None):
        links = self.links if links is None else links
        arr = (abi.LinkDesc * len(links))()
        for i, l in enumerate(links):
            d = arr[i]
            d.parent = l["parent"]
            d.joint = l["joint"]
            d.axis = abi.D3(*l["axis"])
            d.damping = l["damping"]
            d.armature = l["armature"]
            d.free_damping = l["free_damping"]
            rng = l.get("range")
            d.has_range = 1 if rng is not None else 0
            if rng is not None:
                d.range = (abi.C.c_double * 2)(*rng)
            sr, si = l.get("limit_solref"), l.get("limit_solimp")
            lk, lc = l.get("limit_stiffness"), l.get("limit_damping")
            if sr is not None or si is not None or lk is not None or lc is not None:
                sr = sr if sr is not None else (0.02, 1.0)
                si = si if si is not None else (0.9, 0.95, 0.001, 0.5, 2)
                d.limit_custom = 1
                d.limit_solref = (abi.C.c_double * 2)(*sr)
                d.limit_solimp = (abi.C.c_double * 4)(*si[:4])
                d.limit_solimp_power = int(si[4])
                d.limit_stiffness = 1000.0 if lk is None else lk   # JointLimit::DEFAULT
                d.limit_damping = 60.0 if lc is None else lc
            d.parent_pos = abi.D3(*l["parent_pos"])
            d.parent_quat = abi.D4(*l["parent_quat"])
            d.child_pos = abi.D3(*l["child_pos"])
            d.child_quat = abi.D4(*l["child_quat"])
            d.mass = l["mass"]
            d.inertia = (abi.C.c_double * 9)(*l["inertia"])
        return arr

    def geom_structs(self):
        out = []
        for g in self.geoms:
            d = abi.GeomDesc()
            d.shape = g["shape"]
            d.tree = g["tree"]
            d.link = g["link"]
            d.condim = g["condim"]
            d.size = abi.D3(*g["size"])
            d.pos = abi.D3(*g["pos"])
            d.quat = abi.D4(*g["quat"])
            d.friction = g["friction"]
            d.margin = g["margin"]
            d.gap = g["gap"]
            d.solref = (abi.C.c_double * 2)(*g["solref"])
            d.solimp = (abi.C.c_double * 4)(*g["solimp"])
            d.solimp_power = g["power"]
            d.group = g.get("group", 0xFFFFFFFF)
            d.mask = g.get("mask", 0xFFFFFFFF)
            d.body = g.get("body", -1)
            d.mesh = g.get("mesh", -1)
            out.append(d)
        return out

    def sensor_structs(self):
        out = []
        for t in self.sensors:
            d = abi.SensorDesc()
            d.kind = t["kind"]
            d.tree = t.get("tree", 0)
            d.link = t.get("link", -1)
            d.body = t.get("body", -1)
            d.geom = t.get("geom", -1)
            d.offset = abi.D3(*t.get("offset", (0.0, 0.0, 0.0)))
            d.quat = abi.D4(*t.get("quat", (0.0, 0.0, 0.0, 1.0)))
            out.append(d)
        return out

    @property
    def sensor_dim(self):
        return sum(abi.sensor_dim(t["kind"]) for t in self.sensors)

    def body_structs(self):
        out = []
        for b in self.bodies:
            d = abi.BodyDesc()
            d.mass = b["mass"]
            d.inertia = (abi.C.c_double * 9)(*b["inertia"])
            d.position = abi.D3(*b["position"])
            d.orientation = abi.D4(*b["orientation"])
            d.linear_velocity = abi.D3(*b["linear_velocity"])
            d.angular_velocity_body = abi.D3(*b["angular_velocity_body"])
            ml, ma = b.get("max_linear_velocity"), b.get("max_angular_velocity")
            d.has_max_linear_velocity = 0 if ml is None else 1
            d.has_max_angular_velocity = 0 if ma is None else 1
            d.max_linear_velocity = 0.0 if ml is None else ml
            d.max_angular_velocity = 0.0 if ma is None else ma
            d.gravity_scale = b.get("gravity_scale", 1.0)
            out.append(d)
        return out


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ---------------------------------------------------------------------------
# Workload 1: 3-link planar revolute pendulum, no contacts.
# ---------------------------------------------------------------------------
def pendulum(dt=0.002) -> Scene:
    """Fixed root + 3 hinges about +X; capsule links 0.3 m (anchor to anchor),
    radius 0.02 m, 1 kg, COM at mid-link; self collision off, no plane."""
    half = 0.15
    ixx, izz = solid_capsule_inertia(1.0, 0.02, half)
    links = [link(-1, abi.JOINT_FIXED)]
    for i in range(3):
        links.append(link(i, abi.JOINT_HINGE, axis=(1.0, 0.0, 0.0),
                          parent_pos=(0.0, 0.0, 0.0) if i == 0 else (0.0, 0.0, -half),
                          child_pos=(0.0, 0.0, half), mass=1.0, inertia=_diag(ixx, ixx, izz)))
    geoms = [geom(abi.SHAPE_CAPSULE, tree=0, link=i + 1, size=(0.02, half, 0.0)) for i in range(3)]

    def sampler(n, seed):
        r = _rng(seed)
        qpos = r.uniform(-math.pi, math.pi, size=(n, 3))
        qvel = r.normal(0.0, 0.5, size=(n, 3))
        return qpos, qvel

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER,
                solver=abi.SOLVER_PGS, iterations=20)
    return Scene("pendulum3", opts, links, geoms, 0, sampler,
                 "3-link planar revolute pendulum, capsule links 0.3 m x 0.02 m, 1 kg, Euler dt 2 ms")


# ---------------------------------------------------------------------------
# Workload 2: 18-DoF floating-base quadruped on a plane.
# ---------------------------------------------------------------------------
TORSO_HALF = (0.25, 0.15, 0.05)
LIMB_R = 0.03
LIMB_HALF = 0.1   # capsule cylinder half length: 0.2 m limbs


def quadruped(dt=0.002, iterations=50, friction=0.8) -> Scene:
    """Free-root 12 kg box torso and four legs of (hip yaw about Z, hip
    pitch about X, knee about X) hinges; every hinge link is a 0.2 m capsule
    of radius 0.03 m and 1 kg.  The hip and thigh capsules point sideways
    from the torso's long edges, the shank hangs down, so the zero pose
    stands on the four shanks.  Self collision is off (the MJCF loader's
    convention); the ground plane has friction 0.8."""
    links = [link(-1, abi.JOINT_FREE, parent_pos=(0.0, 0.0, 0.4), mass=12.0,
                  inertia=solid_box_inertia(12.0, TORSO_HALF))]
    geoms = [geom(abi.SHAPE_PLANE, friction=friction)]
    geoms.append(geom(abi.SHAPE_BOX, tree=0, link=0, size=TORSO_HALF, friction=1.0))
    ixx, izz = solid_capsule_inertia(1.0, LIMB_R, LIMB_HALF)
    along_y = _diag(ixx, izz, ixx)      # capsule axis along the link's Y
    along_z = _diag(ixx, ixx, izz)      # capsule axis along the link's Z
    to_y = quat_axis_angle((1.0, 0.0, 0.0), -math.pi / 2)   # local Z -> Y
    for sx in (1.0, -1.0):
        for sy in (1.0, -1.0):
            hip = len(links)
            links.append(link(0, abi.JOINT_HINGE, axis=(0.0, 0.0, 1.0),
                              parent_pos=(sx * TORSO_HALF[0], sy * TORSO_HALF[1], 0.0),
                              child_pos=(0.0, -sy * LIMB_HALF, 0.0), mass=1.0, inertia=along_y))
            thigh = len(links)
            links.append(link(hip, abi.JOINT_HINGE, axis=(1.0, 0.0, 0.0),
                              parent_pos=(0.0, sy * LIMB_HALF, 0.0),
                              child_pos=(0.0, -sy * LIMB_HALF, 0.0), mass=1.0, inertia=along_y))
            shank = len(links)
            links.append(link(thigh, abi.JOINT_HINGE, axis=(1.0, 0.0, 0.0),
                              parent_pos=(0.0, sy * LIMB_HALF, 0.0),
                              child_pos=(0.0, 0.0, LIMB_HALF), mass=1.0, inertia=along_z))
            geoms.append(geom(abi.SHAPE_CAPSULE, tree=0, link=hip, size=(LIMB_R, LIMB_HALF, 0.0), quat=to_y))
            geoms.append(geom(abi.SHAPE_CAPSULE, tree=0, link=thigh, size=(LIMB_R, LIMB_HALF, 0.0), quat=to_y))
            geoms.append(geom(abi.SHAPE_CAPSULE, tree=0, link=shank, size=(LIMB_R, LIMB_HALF, 0.0)))

    def sampler(n, seed):
        r = _rng(seed)
        qpos = np.zeros((n, 19))
        qpos[:, 2] = r.uniform(0.3, 0.5, size=n)
        qpos[:, 6] = 1.0
        qpos[:, 7:] = r.normal(0.0, 0.2, size=(n, 12))
        qvel = np.zeros((n, 18))
        return qpos, qvel

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER,
                solver=abi.SOLVER_PGS, iterations=iterations)
    return Scene("quadruped18", opts, links, geoms, 0, sampler,
                 "18-DoF floating-base quadruped on a plane (mu 0.8), PGS 50, Euler dt 2 ms")


# ---------------------------------------------------------------------------
# Large contact sets: a chain of boxes lying on the plane.
# ---------------------------------------------------------------------------
BOX_HALF = (0.1, 0.05, 0.03)


def box_chain(n_boxes=6, dt=0.002, iterations=30) -> Scene:
    """Free-root box and n_boxes-1 further boxes joined end to end by hinges
    about the vertical axis, all lying on the plane: up to 4 corner contacts
    per box (n_boxes=6: 24 contacts, 96 pyramidal rows), past the 64 rows a
    wavefront holds in registers.  1 kg boxes, friction 1.0."""
    hx = BOX_HALF[0]
    inertia = solid_box_inertia(1.0, BOX_HALF)
    links = [link(-1, abi.JOINT_FREE, parent_pos=(0.0, 0.0, BOX_HALF[2]), mass=1.0, inertia=inertia)]
    for i in range(1, n_boxes):
        links.append(link(i - 1, abi.JOINT_HINGE, axis=(0.0, 0.0, 1.0), parent_pos=(hx + 0.01, 0.0, 0.0),
                          child_pos=(-hx - 0.01, 0.0, 0.0), mass=1.0, inertia=inertia))
    geoms = [geom(abi.SHAPE_PLANE, friction=1.0)]
    for i in range(n_boxes):
        geoms.append(geom(abi.SHAPE_BOX, tree=0, link=i, size=BOX_HALF, friction=1.0))

    def sampler(n, seed):
        r = _rng(seed)
        qpos = np.zeros((n, 7 + n_boxes - 1))
        qpos[:, 0:2] = r.normal(0.0, 0.05, size=(n, 2))
        qpos[:, 2] = BOX_HALF[2] + r.uniform(0.0, 0.01, size=n)
        qpos[:, 6] = 1.0
        qpos[:, 7:] = r.normal(0.0, 0.3, size=(n, n_boxes - 1))
        qvel = np.zeros((n, 6 + n_boxes - 1))
        qvel[:, 3:5] = r.normal(0.0, 0.2, size=(n, 2))
        qvel[:, 6:] = r.normal(0.0, 0.5, size=(n, n_boxes - 1))
        return qpos, qvel

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER,
                solver=abi.SOLVER_PGS, iterations=iterations)
    return Scene(f"boxchain{n_boxes}", opts, links, geoms, 0, sampler,
                 f"{n_boxes} hinged boxes lying on a plane (up to {4 * n_boxes} contacts), PGS {iterations}")


# ---------------------------------------------------------------------------
# Workload 4: fixed-root hinge chains of 2..64 links draped on a plane.
# ---------------------------------------------------------------------------
CHAIN_LINK = 0.1     # anchor-to-anchor link length
CHAIN_R = 0.02
CHAIN_ARMATURE = 0.01  # joint armature: without it the longest chains diverge at dt 1 ms (in newt's step too)


def hinge_chain(n, dt=0.001, iterations=30, seed=2026) -> Scene:
    """Fixed root and n hinge links (axes alternating X and Y), each a
    capsule of radius 0.02 m spanning its 0.1 m link, masses ~ U[0.5, 2] kg
    (seeded per chain length), joint armature 0.01 kg m².  With k = max(4, n // 4 + 2) the root hangs
    max(n - k, 0) link lengths (plus a radius and 5 mm) above the plane and
    the sampled states bend the chain by a right angle at that link, so the
    last (up to) k links start just above the plane and settle onto it."""
    rm = _rng(seed * 131 + n)
    masses = rm.uniform(0.5, 2.0, size=n)
    half = CHAIN_LINK / 2
    bend = max(n - max(4, n // 4 + 2), 0)
    height = bend * CHAIN_LINK + CHAIN_R + 0.005
    links = [link(-1, abi.JOINT_FIXED, parent_pos=(0.0, 0.0, height))]
    geoms = [geom(abi.SHAPE_PLANE, friction=0.8)]
    for i in range(n):
        ixx, izz = solid_capsule_inertia(masses[i], CHAIN_R, half - CHAIN_R)
        axis = (1.0, 0.0, 0.0) if i % 2 == 0 else (0.0, 1.0, 0.0)
        links.append(link(i, abi.JOINT_HINGE, axis=axis,
                          parent_pos=(0.0, 0.0, 0.0) if i == 0 else (0.0, 0.0, -half),
                          child_pos=(0.0, 0.0, half), mass=float(masses[i]), inertia=_diag(ixx, ixx, izz),
                          armature=CHAIN_ARMATURE))
        geoms.append(geom(abi.SHAPE_CAPSULE, tree=0, link=i + 1, size=(CHAIN_R, half - CHAIN_R, 0.0), friction=0.8))

    def sampler(nw, seed_):
        r = _rng(seed_)
        q = r.normal(0.0, 0.05, size=(nw, n))
        q[:, bend] += math.pi / 2
        return q, np.zeros((nw, n))

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER,
                solver=abi.SOLVER_PGS, iterations=iterations)
    return Scene(f"chain{n}", opts, links, geoms, 0, sampler,
                 f"fixed-root {n}-hinge capsule chain draped on a plane, PGS {iterations}, Euler dt {dt}")


def hinge_coil(n=10, dt=0.001, iterations=30) -> Scene:
    """Planar chain of n hinges (axes X) under a fixed root, folded like an
    accordion (bends of about ±2.6 rad) so that links two apart lie close
    together; capsule links (0.1 m, radius 0.02 m, 0.5 kg) and a sphere of
    radius 0.025 m at the tip of every even link.  Self collision is on;
    collision groups let geoms of links of equal parity collide with each
    other (never neighbours) and every geom with the plane below, giving
    capsule-capsule, sphere-capsule, sphere-sphere and plane contacts."""
    half = CHAIN_LINK / 2
    height = 0.35
    links = [link(-1, abi.JOINT_FIXED, parent_pos=(0.0, 0.0, height))]
    plane_bit = 4
    geoms = [geom(abi.SHAPE_PLANE, friction=0.8, group=7, mask=7)]
    ixx, izz = solid_capsule_inertia(0.5, CHAIN_R, half - CHAIN_R)
    for i in range(n):
        links.append(link(i, abi.JOINT_HINGE, axis=(1.0, 0.0, 0.0),
                          parent_pos=(0.0, 0.0, 0.0) if i == 0 else (0.0, 0.0, -half),
                          child_pos=(0.0, 0.0, half), mass=0.5, inertia=_diag(ixx, ixx, izz),
                          armature=CHAIN_ARMATURE))
        grp = 1 << (i % 2)
        geoms.append(geom(abi.SHAPE_CAPSULE, tree=0, link=i + 1, size=(CHAIN_R, half - CHAIN_R, 0.0), friction=0.8,
                          group=grp, mask=grp | plane_bit))
        if i % 2 == 0:
            geoms.append(geom(abi.SHAPE_SPHERE, tree=0, link=i + 1, size=(0.025, 0.0, 0.0), pos=(0.0, 0.0, -half),
                              friction=0.8, group=grp, mask=grp | plane_bit))

    def sampler(nw, seed_):
        r = _rng(seed_)
        sign = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
        q = sign * 2.6 + r.normal(0.0, 0.15, size=(nw, n))
        q[:, 0] = r.normal(0.0, 0.3, size=nw)
        return q, r.normal(0.0, 0.5, size=(nw, n))

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER,
                solver=abi.SOLVER_PGS, iterations=iterations)
    return Scene(f"coil{n}", opts, links, geoms, 1, sampler,
                 f"folded {n}-hinge capsule chain with self collision, PGS {iterations}, Euler dt {dt}")


def box_corridor(n_boxes=6, dt=0.002, iterations=30) -> Scene:
    """The hinged box chain of box_chain() between two static box walls
    (half extents 1.0 x 0.03 x 0.1 m, 4 mm beyond the boxes' sides) that the
    swinging boxes strike, with self collision between boxes two or more
    links apart (collision groups): box-box contacts with one side on a link
    and the other static or on another link, beside box-plane contacts."""
    base = box_chain(n_boxes, dt, iterations)
    geoms = [dict(base.geoms[0], group=7, mask=7)]
    for i, g in enumerate(base.geoms[1:]):
        grp = 1 << (i % 2)
        geoms.append(dict(g, group=grp, mask=grp | 4))
    wall_y = BOX_HALF[1] + 0.03 + 0.004
    cx = (n_boxes - 1) * (BOX_HALF[0] + 0.01)
    for sgn in (1.0, -1.0):
        geoms.append(geom(abi.SHAPE_BOX, size=(1.0, 0.03, 0.1), pos=(cx, sgn * wall_y, 0.1), friction=0.6,
                          group=7, mask=7))

    def sampler(n, seed):
        q, v = base.sampler(n, seed)
        r = _rng(seed + 1)
        q[:, 7:] *= 0.3
        v[:, 3:5] = 0.0
        v[:, 4] = r.normal(0.0, 0.3, size=n)
        v[:, 6:] = r.normal(0.0, 1.5, size=(n, n_boxes - 1))
        return q, v

    return Scene(f"boxcorridor{n_boxes}", dict(base.options), base.links, geoms, 1, sampler,
                 f"{n_boxes} hinged boxes between two static walls, self collision on, PGS {iterations}")


def with_options(sc: Scene, suffix: str, **options) -> Scene:
    """A copy of `sc` with World options replaced, e.g.
    ``with_options(pendulum(), "_rk4", integrator=abi.INTEGRATOR_RK4)``."""
    o = dict(sc.options)
    o.update(options)
    return Scene(sc.name + suffix, o, sc.links, sc.geoms, sc.self_collide, sc.sampler,
                 f"{sc.description}; options {options}", dict(sc.extra), list(sc.bodies), list(sc.sensors),
                 sc.magnetic_field, list(sc.more_trees), list(sc.meshes))


def _body_state(bodies):
    q = np.array([[*b["position"], *b["orientation"]] for b in bodies], dtype=np.float64).reshape(-1)
    v = np.array([[*b["linear_velocity"], *b["angular_velocity_body"]] for b in bodies], dtype=np.float64).reshape(-1)
    return q, v


def box_stack(n_boxes=3, dt=0.005, iterations=30, integrator=abi.INTEGRATOR_RK4) -> Scene:
    """newt's solver golden stack (newt/tests/solver_golden.rs:27-46): boxes of
    half extent 0.25 and 1 kg stacked on a plane with 1 cm x offsets per
    level, friction 0.8, SolRef (0.05, 1.5); PGS, pyramidal cone.  Worlds
    differ by a small random lateral velocity per box."""
    half = (0.25, 0.25, 0.25)
    bodies = [body(1.0, solid_box_inertia(1.0, half), (0.01 * k, 0.0, 0.25 + 0.5 * k)) for k in range(n_boxes)]
    geoms = [geom(abi.SHAPE_PLANE, friction=0.8, solref=(0.05, 1.5))]
    geoms += [geom(abi.SHAPE_BOX, body=k, size=half, friction=0.8, solref=(0.05, 1.5)) for k in range(n_boxes)]
    q0, v0 = _body_state(bodies)

    def sampler(n, seed):
        r = _rng(seed)
        q = np.tile(q0, (n, 1))
        v = np.tile(v0, (n, 1))
        for k in range(n_boxes):
            v[:, 6 * k:6 * k + 2] = r.normal(0.0, 0.05, size=(n, 2))
        return q, v

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=integrator, solver=abi.SOLVER_PGS,
                iterations=iterations)
    return Scene(f"boxstack{n_boxes}", opts, [], geoms, 0, sampler,
                 f"{n_boxes} stacked free boxes on a plane, PGS {iterations}", bodies=bodies)


def _free_sphere_link(z=0.0, x=0.0):
    """newt/tests/tree_contacts_solver.rs:17-28: a one-link tree on a free
    joint at (x, 0, z), 1 kg, inertia diag 0.1."""
    return link(-1, abi.JOINT_FREE, parent_pos=(x, 0.0, z), mass=1.0, inertia=_diag(0.1, 0.1, 0.1))


def tree_body_spheres(solver=abi.SOLVER_PGS, as_tree=False) -> Scene:
    """tree_body_contact_transfers_equal_and_opposite_impulses
    (newt/tests/tree_contacts_solver.rs:404-448): a free-root sphere tree at
    x = -0.4 and a 2 kg free sphere body at x = +0.4, radius 0.5 each and
    frictionless, overlapping by 0.2 m at rest; no gravity, 30 iterations,
    World::new's dt 5 ms and RK4.  With `as_tree` the body is a second
    free-root tree of the same mass and inertia instead."""
    links = [_free_sphere_link(0.0, -0.4)]
    inertia = _diag(0.4 * 2.0 * 0.25, 0.4 * 2.0 * 0.25, 0.4 * 2.0 * 0.25)
    geoms = [geom(abi.SHAPE_SPHERE, tree=0, link=0, size=(0.5, 0.0, 0.0), friction=0.0)]
    bodies, more = [], []
    if as_tree:
        more = [([link(-1, abi.JOINT_FREE, parent_pos=(0.4, 0.0, 0.0), mass=2.0, inertia=inertia)], 0)]
        geoms.append(geom(abi.SHAPE_SPHERE, tree=1, link=0, size=(0.5, 0.0, 0.0), friction=0.0))
    else:
        bodies = [body(2.0, inertia, (0.4, 0.0, 0.0))]
        geoms.append(geom(abi.SHAPE_SPHERE, body=0, size=(0.5, 0.0, 0.0), friction=0.0))
    opts = dict(dt=0.005, gravity=(0.0, 0.0, 0.0), integrator=abi.INTEGRATOR_RK4, solver=solver, iterations=30)
    sc = Scene("tree_body_spheres" + ("_trees" if as_tree else ""), opts, links, geoms, 0, None,
               "a free sphere tree and a free sphere body overlapping at rest", bodies=bodies, more_trees=more)

    def sampler(n, seed):
        q = np.zeros((n, sc.nq))
        q[:, 0] = -0.4
        q[:, 6] = 1.0
        q[:, 7] = 0.4
        q[:, 13] = 1.0
        return q, np.zeros((n, sc.nv))

    sc.sampler = sampler
    return sc


def tree_body_tilted_boxes(solver=abi.SOLVER_PGS, integrator=None, as_tree=False) -> Scene:
    """tilted_tree_box_box_world (newt/tests/tree_contacts_solver.rs:105-143):
    a free-root tree at z 1.48 tilted 0.005 rad about x carrying a box of
    half extent 0.5, over a 1 kg free box body at z 0.5 tilted 0.01 rad;
    frictionless, no gravity, 40 iterations, dt 5 ms, RK4 unless
    `integrator` says otherwise.  With `as_tree` the body is a second
    free-root tree of the same mass and inertia."""
    half = (0.5, 0.5, 0.5)
    q_tree = quat_axis_angle((1.0, 0.0, 0.0), 0.005)
    q_body = quat_axis_angle((1.0, 0.0, 0.0), 0.01)
    links = [_free_sphere_link(1.48)]
    geoms = [geom(abi.SHAPE_BOX, tree=0, link=0, size=half, friction=0.0)]
    bodies, more = [], []
    if as_tree:
        more = [([link(-1, abi.JOINT_FREE, parent_pos=(0.0, 0.0, 0.5), mass=1.0,
                       inertia=solid_box_inertia(1.0, half))], 0)]
        geoms.append(geom(abi.SHAPE_BOX, tree=1, link=0, size=half, friction=0.0))
    else:
        bodies = [body(1.0, solid_box_inertia(1.0, half), (0.0, 0.0, 0.5), q_body)]
        geoms.append(geom(abi.SHAPE_BOX, body=0, size=half, friction=0.0))
    opts = dict(dt=0.005, gravity=(0.0, 0.0, 0.0),
                integrator=abi.INTEGRATOR_RK4 if integrator is None else integrator, solver=solver, iterations=40)
    sc = Scene("tree_body_tilted_boxes" + ("_trees" if as_tree else ""), opts, links, geoms, 0, None,
               "a tilted tree box resting on a tilted box body", bodies=bodies, more_trees=more)

    def sampler(n, seed):
        q = np.zeros((n, sc.nq))
        q[:, 2] = 1.48
        q[:, 3:7] = q_tree
        q[:, 9] = 0.5
        q[:, 10:14] = q_body
        return q, np.zeros((n, sc.nv))

    sc.sampler = sampler
    return sc


def chain_and_boxes(n_links=4, n_free=4, solver=abi.SOLVER_PGS, integrator=abi.INTEGRATOR_EULER, dt=0.002,
                    iterations=30) -> Scene:
    """A tree and free bodies in one world: the hinged box chain of
    box_chain(n_links) lying on a plane with n_free free boxes (half extents
    0.06 x 0.05 x 0.04, 0.5 kg) dropped onto its links from 0.15 m up, then
    onto each other and the plane: tree-body, body-body, body-plane and
    tree-plane contacts.  Worlds differ by the chain's pose and joint
    angles (box_chain's sampler) and each box's lateral offset and
    orientation."""
    base = box_chain(n_links, dt=dt, iterations=iterations)
    h = (0.06, 0.05, 0.04)
    bodies = []
    geoms = list(base.geoms)
    for i in range(n_free):
        x = (i % n_links - (n_links - 1) / 2.0) * 0.08
        bodies.append(body(0.5, solid_box_inertia(0.5, h), (x, 0.0, 0.15 + 0.1 * i)))
        geoms.append(geom(abi.SHAPE_BOX, body=i, size=h, friction=1.0))
    opts = dict(base.options, solver=solver, integrator=integrator)
    sc = Scene(f"chain{n_links}_boxes{n_free}", opts, base.links, geoms, 0, None,
               f"{n_links} hinged boxes on a plane with {n_free} free boxes dropped on them", bodies=bodies)

    def sampler(n, seed):
        q, v = base.sampler(n, seed)
        r = _rng(seed + 1)
        bq = np.zeros((n, 7 * n_free))
        for i in range(n_free):
            b = bodies[i]
            bq[:, 7 * i:7 * i + 3] = np.array(b["position"]) + np.c_[r.uniform(-0.03, 0.03, size=(n, 2)),
                                                                     np.zeros(n)]
            bq[:, 7 * i + 3:7 * i + 7] = random_quaternions(r, n)
        return np.hstack([q, bq]), np.hstack([v, np.zeros((n, 6 * n_free))])

    sc.sampler = sampler
    return sc


def box_field(n_x=10, n_y=5, levels=2, half=0.1, spacing=0.35, dt=0.002, iterations=20, yaw=0.2) -> Scene:
    """n_x x n_y stacks of `levels` cubes (half extent `half`, 1 kg) resting
    on a plane 0.1 mm into contact, stacks `spacing` apart; Newton, pyramidal,
    Euler.  Every cube rests on a face, so a world holds about 4 contacts per
    cube (at the defaults 100 cubes, about 400 contacts and 1600 Newton
    rows).  Worlds differ by a random yaw (up to `yaw` radians) and a
    lateral offset (up to 5 mm) per cube."""
    bodies, geoms = [], [geom(abi.SHAPE_PLANE, friction=0.8)]
    hs = (half, half, half)
    for i in range(n_x):
        for j in range(n_y):
            for k in range(levels):
                p = ((i - (n_x - 1) / 2) * spacing, (j - (n_y - 1) / 2) * spacing, half - 1e-4 + k * (2 * half - 1e-4))
                geoms.append(geom(abi.SHAPE_BOX, body=len(bodies), size=hs, friction=0.8))
                bodies.append(body(1.0, solid_box_inertia(1.0, hs), p))
    q0, v0 = _body_state(bodies)
    nb = len(bodies)

    def sampler(n, seed):
        r = _rng(seed)
        q = np.tile(q0, (n, 1))
        v = np.tile(v0, (n, 1))
        for b in range(nb):
            q[:, 7 * b:7 * b + 2] += r.uniform(-0.005, 0.005, size=(n, 2))
            z = r.uniform(-yaw, yaw, size=n)
            q[:, 7 * b + 3:7 * b + 7] = np.stack([np.zeros(n), np.zeros(n), np.sin(z / 2), np.cos(z / 2)], axis=1)
        return q, v

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER, solver=abi.SOLVER_NEWTON,
                iterations=iterations)
    return Scene(f"boxfield{nb}", opts, [], geoms, 0, sampler,
                 f"{n_x}x{n_y} stacks of {levels} cubes resting on a plane, Newton {iterations}, Euler dt {dt}",
                 bodies=bodies)


def penalty_stack(dt=0.005, integrator=abi.INTEGRATOR_RK4) -> Scene:
    """newt's penalty contact golden scene (newt/tests/contacts_golden.rs:31-59):
    three 1 kg boxes of half extent 0.35 dropped onto a plane (friction 0.6,
    default SolRef), the middle one 2 cm off in x, the top one spinning at
    0.3 rad/s about its y axis; the penalty solver (spring-damper contacts,
    legacy box-box manifold).  Worlds differ by a small random lateral
    velocity per box."""
    half = (0.35, 0.35, 0.35)
    drops = [((0.0, 0.0, 0.5), (0.0, 0.0, 0.0)), ((0.02, 0.0, 1.7), (0.0, 0.0, 0.0)),
             ((0.0, 0.0, 2.9), (0.0, 0.3, 0.0))]
    bodies = [body(1.0, solid_box_inertia(1.0, half), p, angular_velocity_body=w) for p, w in drops]
    geoms = [geom(abi.SHAPE_PLANE, friction=0.6)]
    geoms += [geom(abi.SHAPE_BOX, body=k, size=half, friction=0.6) for k in range(3)]
    q0, v0 = _body_state(bodies)

    def sampler(n, seed):
        r = _rng(seed)
        q = np.tile(q0, (n, 1))
        v = np.tile(v0, (n, 1))
        for k in range(3):
            v[:, 6 * k:6 * k + 2] = r.normal(0.0, 0.05, size=(n, 2))
        return q, v

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=integrator, solver=abi.SOLVER_PENALTY, iterations=20)
    return Scene("penaltystack3", opts, [], geoms, 0, sampler, "3 boxes dropped on a plane, penalty contacts",
                 bodies=bodies)


def random_quaternions(r, n):
    """Uniform random unit quaternions (x, y, z, w), Shoemake's method."""
    u1, u2, u3 = r.random(n), r.random(n), r.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3),
                     b * np.cos(2 * np.pi * u3)], axis=1)


def box_pile(n_side=4, dt=0.002, iterations=20, seed=2026, friction=0.8) -> Scene:
    """n_side^3 free boxes dropped onto a plane: half extents ~ U[0.05, 0.2]^3
    (one draw, shared by every world), density 1000 (mass = 8 hx hy hz 1000),
    a jittered n_side^3 grid at 0.5 m spacing from 0.35 m up, uniform random
    orientations per world; PGS, pyramidal cone, Euler.  At n_side 4: 64
    boxes, 65 geoms, 2080 candidate pairs."""
    r = _rng(seed)
    nb = n_side ** 3
    halves = r.uniform(0.05, 0.2, size=(nb, 3))
    bodies, geoms = [], [geom(abi.SHAPE_PLANE, friction=friction)]
    grid = []
    for i in range(n_side):
        for j in range(n_side):
            for k in range(n_side):
                grid.append(((i - (n_side - 1) / 2) * 0.5, (j - (n_side - 1) / 2) * 0.5, 0.35 + 0.5 * k))
    for b in range(nb):
        h = tuple(float(x) for x in halves[b])
        mass = 8.0 * h[0] * h[1] * h[2] * 1000.0
        bodies.append(body(mass, solid_box_inertia(mass, h), grid[b]))
        geoms.append(geom(abi.SHAPE_BOX, body=b, size=h, friction=friction))
    grid = np.array(grid)

    def sampler(n, s):
        rr = _rng(s)
        q = np.zeros((n, 7 * nb))
        v = np.zeros((n, 6 * nb))
        for b in range(nb):
            q[:, 7 * b:7 * b + 3] = grid[b] + rr.uniform(-0.04, 0.04, size=(n, 3))
            q[:, 7 * b + 3:7 * b + 7] = random_quaternions(rr, n)
        return q, v

    q0, _ = sampler(1, seed)
    for b in range(nb):
        bodies[b]["position"] = tuple(q0[0, 7 * b:7 * b + 3])
        bodies[b]["orientation"] = tuple(q0[0, 7 * b + 3:7 * b + 7])
    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER, solver=abi.SOLVER_PGS,
                iterations=iterations)
    return Scene(f"boxpile{nb}", opts, [], geoms, 0, sampler,
                 f"{nb} free boxes dropped on a plane, PGS {iterations}, Euler dt {dt}", bodies=bodies)


# ---------------------------------------------------------------------------
# Convex meshes (GeomShape::Mesh): GJK/EPA mesh-mesh, mesh-plane and
# sphere-mesh contacts.
# ---------------------------------------------------------------------------
TETRA_MESH = (np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),
              np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], dtype=np.int32))


def cube_mesh(h):
    """A cube of half extent h as 8 vertices and 12 outward-wound triangles."""
    v = np.array([[sx * h, sy * h, sz * h] for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)], dtype=np.float64)
    f = np.array([[0, 3, 1], [0, 2, 3], [4, 5, 7], [4, 7, 6], [0, 1, 5], [0, 5, 4],
                  [2, 6, 7], [2, 7, 3], [0, 4, 6], [0, 6, 2], [1, 3, 7], [1, 7, 5]], dtype=np.int32)
    return v, f


def mesh_tumble(rotated=False) -> Scene:
    """newt's dynamic mesh anchor world (newt/tests/contacts_geoms_v1.rs:
    2194-2229): a static unit tetrahedron and a free one (mass 1, principal
    inertia 0.166667) overlapping it and spinning, zero gravity, dt 5 ms,
    World::new's RK4 and penalty contacts."""
    if rotated:
        pos, q, w = (0.25, 0.1, 0.25), quat_axis_angle((0.0, 0.0, 1.0), math.radians(0.4)), (-0.6, 0.9, 0.5)
    else:
        pos, q, w = (0.2, 0.2, 0.2), IDENTITY_Q, (1.0, 0.7, -0.4)
    i = 0.166667
    b = body(1.0, _diag(i, i, i), pos, q, angular_velocity_body=w)
    geoms = [geom(abi.SHAPE_MESH, mesh=0, friction=0.5), geom(abi.SHAPE_MESH, body=0, mesh=0, friction=0.5)]
    qpos = np.array([list(pos) + list(q)])
    qvel = np.array([[0.0, 0.0, 0.0] + list(w)])
    opts = dict(dt=0.005, gravity=(0.0, 0.0, 0.0), integrator=abi.INTEGRATOR_RK4, solver=abi.SOLVER_PENALTY,
                iterations=20)
    return Scene("mesh_tumble" + ("_rotated" if rotated else ""), opts, [], geoms, 0,
                 lambda n, seed: (np.repeat(qpos, n, 0), np.repeat(qvel, n, 0)),
                 "a tetrahedron tumbling inside another (penalty, RK4)", bodies=[b], meshes=[TETRA_MESH])


def mesh_pile(dt=0.002, iterations=30, seed=7, friction=0.8) -> Scene:
    """Free convex meshes dropped onto a plane: three cubes (half extent 0.1),
    three tetrahedra (edge 0.3) and a sphere (r 0.08) in a jittered stack
    0.33-0.4 m apart, uniform random orientations per world; PGS, pyramidal cone,
    Euler.  Mesh-mesh (GJK/EPA and the face manifold), mesh-plane and
    sphere-mesh contacts."""
    cube = cube_mesh(0.1)
    tet = (TETRA_MESH[0] * 0.3 - 0.075, TETRA_MESH[1])
    kinds = ["cube", "tet", "cube", "tet", "sphere", "cube", "tet"]
    bodies, geoms = [], [geom(abi.SHAPE_PLANE, friction=friction)]
    spots = [(0.0, 0.0, 0.15), (0.05, 0.02, 0.55), (-0.04, 0.06, 0.95), (0.02, -0.05, 1.35), (0.0, 0.03, 1.72),
             (0.06, 0.0, 2.05), (-0.03, -0.02, 2.45)]
    for b, k in enumerate(kinds):
        if k == "cube":
            m = 8.0 * 0.1 ** 3 * 500.0
            bodies.append(body(m, solid_box_inertia(m, (0.1, 0.1, 0.1)), spots[b]))
            geoms.append(geom(abi.SHAPE_MESH, body=b, mesh=0, friction=friction))
        elif k == "tet":
            m = 2.0
            bodies.append(body(m, solid_box_inertia(m, (0.08, 0.08, 0.08)), spots[b]))
            geoms.append(geom(abi.SHAPE_MESH, body=b, mesh=1, friction=friction))
        else:
            m = 1.0
            ii = 0.4 * m * 0.08 ** 2
            bodies.append(body(m, _diag(ii, ii, ii), spots[b]))
            geoms.append(geom(abi.SHAPE_SPHERE, body=b, size=(0.08, 0.0, 0.0), friction=friction))
    nb = len(kinds)
    grid = np.array(spots)

    def sampler(n, s):
        rr = _rng(s)
        q = np.zeros((n, 7 * nb))
        v = np.zeros((n, 6 * nb))
        for b in range(nb):
            q[:, 7 * b:7 * b + 3] = grid[b] + rr.uniform(-0.03, 0.03, size=(n, 3))
            q[:, 7 * b + 3:7 * b + 7] = random_quaternions(rr, n)
        return q, v

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER, solver=abi.SOLVER_PGS,
                iterations=iterations)
    return Scene("meshpile7", opts, [], geoms, 0, sampler,
                 "cube and tetrahedron meshes and a sphere dropped on a plane, PGS, Euler", bodies=bodies,
                 meshes=[cube, tet])


def mesh_tree(dt=0.002, iterations=30) -> Scene:
    """A free-root tree whose root carries a cube mesh (half extent 0.12) and a
    hinged child a tetrahedron mesh, dropped onto a plane beside a static cube
    mesh obstacle; self collision on.  PGS, RK4.  Tree mesh-plane and
    link-static and link-link mesh-mesh contacts."""
    cube = cube_mesh(0.12)
    tet = (TETRA_MESH[0] * 0.25, TETRA_MESH[1])
    links = [link(-1, abi.JOINT_FREE, mass=3.0, inertia=solid_box_inertia(3.0, (0.12, 0.12, 0.12))),
             link(0, abi.JOINT_HINGE, axis=(0.0, 1.0, 0.0), parent_pos=(0.15, 0.0, -0.05), child_pos=(0.0, 0.0, 0.1),
                  mass=0.8, inertia=solid_box_inertia(0.8, (0.08, 0.08, 0.08)), damping=0.05)]
    geoms = [geom(abi.SHAPE_PLANE, friction=0.9),
             geom(abi.SHAPE_MESH, tree=0, link=0, mesh=0, friction=0.9),
             geom(abi.SHAPE_MESH, tree=0, link=1, mesh=1, pos=(0.0, 0.0, -0.1), friction=0.9),
             geom(abi.SHAPE_MESH, mesh=0, pos=(0.35, 0.0, 0.12), friction=0.9)]

    def sampler(n, s):
        rr = _rng(s)
        q = np.zeros((n, 8))
        q[:, :3] = np.array([0.05, 0.0, 0.5]) + rr.uniform(-0.05, 0.05, size=(n, 3))
        q[:, 3:7] = random_quaternions(rr, n)
        q[:, 7] = rr.uniform(-0.6, 0.6, size=n)
        v = np.zeros((n, 7))
        v[:, :3] = rr.normal(0.0, 0.3, size=(n, 3))
        return q, v

    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_RK4, solver=abi.SOLVER_PGS,
                iterations=iterations)
    return Scene("meshtree", opts, links, geoms, 1, sampler,
                 "a cube-mesh root with a hinged tetrahedron mesh dropped beside a static cube mesh, PGS, RK4",
                 meshes=[cube, tet])


SCENES = {"pendulum3": pendulum, "quadruped18": quadruped, "boxchain6": box_chain, "coil10": hinge_coil,
          "chain40": lambda: hinge_chain(40), "chain64": lambda: hinge_chain(64), "boxstack3": box_stack,
          "boxpile64": box_pile,
          "boxpile64_newton": lambda: with_options(box_pile(), "_newton", solver=abi.SOLVER_NEWTON),
          "chain4_boxes4": chain_and_boxes}


def sensor(kind, link=-1, body=-1, geom=-1, offset=(0.0, 0.0, 0.0), quat=(0.0, 0.0, 0.0, 1.0), tree=0):
    """One sensor (newt's Sensor with its SiteFrame, sensor.rs:118-122,
    226-229): link `link` of tree `tree` or free body `body` is the site's
    parent (or the joint/force sensor's link), `geom` the touch sensor's
    geom."""
    return dict(kind=kind, tree=tree, link=link, body=body, geom=geom, offset=tuple(offset), quat=tuple(quat))
