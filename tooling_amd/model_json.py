"""newt's JSON scene format (newt/src/model.rs, docs/model-format.md) for the
models the GPU step builds.

`load(path_or_text)` returns a :class:`tooling_amd.scenes.Scene` that
:class:`tooling_amd.newt.Model` compiles.  The loader follows newt's
strict-by-default rule: it rejects unknown fields at every scope, and every
error names the JSON path of the offending value (``ModelError``).  It also
rejects what the GPU step does not build, rather than dropping it:
- trees together with free bodies;
- rolling friction on bodies;
- actuators, equalities, tendons and heightfields;
- ball joints, cylinders and ellipsoids;
- explicit contact-pair lists.

Convex mesh assets (top-level ``meshes``, referenced by mesh geoms by name)
load into ``scene.meshes``.  Sites are parsed and kept (``scene.extra["sites"]``); they do not take part
in the step.

Numbers are rounded to float32 as newt parses them (``get_f32``), and derived
quantities (solid-shape inertias, normalized axes) are computed in float32
as newt does, so a loaded model holds newt's values exactly.
"""
from __future__ import annotations

import json
import math
import os
from typing import Any, Dict, List

import numpy as np

from . import abi, scenes


class ModelError(ValueError):
    def __init__(self, path: str, message: str):
        super().__init__(f"{path}: {message}")
        self.path = path


f32 = np.float32


def _obj(v, path, allowed):
    if not isinstance(v, dict):
        raise ModelError(path, "expected an object")
    for k in v:
        if k not in allowed:
            raise ModelError(f"{path}.{k}", "unknown field")
    return v


def _num(v, path, positive=False, nonneg=False):
    if isinstance(v, bool) or not isinstance(v, (int, float)) or not math.isfinite(v):
        raise ModelError(path, "expected a finite number")
    x = f32(v)
    if positive and not x > 0:
        raise ModelError(path, "must be > 0")
    if nonneg and not x >= 0:
        raise ModelError(path, "must be >= 0")
    return x


def _vec(v, n, path):
    if not isinstance(v, list) or len(v) != n:
        raise ModelError(path, f"expected an array of {n} numbers")
    return [_num(x, f"{path}[{i}]") for i, x in enumerate(v)]


def _quat_normalized(q):
    """Quat::renormalize (math.rs:590-597) in float32."""
    q = [f32(x) for x in q]
    n = f32(math.sqrt(f32(f32(f32(q[0] * q[0]) + f32(q[1] * q[1])) + f32(q[2] * q[2])) + f32(q[3] * q[3])))
    if n == 0:
        return [f32(0), f32(0), f32(0), f32(1)]
    return [f32(x / n) for x in q]


def _pose(v, path, allow_orientation=True):
    """{"position": [..], "orientation": [x,y,z,w] | "orientation_axis_angle": {...}}."""
    if v is None:
        return [f32(0)] * 3, [f32(0), f32(0), f32(0), f32(1)]
    _obj(v, path, {"position", "orientation", "orientation_axis_angle"})
    pos = _vec(v.get("position", [0, 0, 0]), 3, f"{path}.position")
    if "orientation" in v and "orientation_axis_angle" in v:
        raise ModelError(path, "give orientation or orientation_axis_angle, not both")
    quat = [f32(0), f32(0), f32(0), f32(1)]
    if "orientation" in v:
        quat = _quat_normalized(_vec(v["orientation"], 4, f"{path}.orientation"))
    elif "orientation_axis_angle" in v:
        aa = _obj(v["orientation_axis_angle"], f"{path}.orientation_axis_angle", {"axis", "angle"})
        axis = _vec(aa.get("axis"), 3, f"{path}.orientation_axis_angle.axis")
        angle = _num(aa.get("angle"), f"{path}.orientation_axis_angle.angle")
        ax = _normalized(axis, f"{path}.orientation_axis_angle.axis")
        s, c = f32(math.sin(f32(angle * f32(0.5)))), f32(math.cos(f32(angle * f32(0.5))))
        quat = [f32(ax[0] * s), f32(ax[1] * s), f32(ax[2] * s), c]
    if not allow_orientation and quat != [0, 0, 0, 1]:
        raise ModelError(path, "non-root joint offsets must have identity orientation")
    return pos, quat


def _normalized(a, path):
    n = f32(math.sqrt(f32(f32(f32(a[0] * a[0]) + f32(a[1] * a[1])) + f32(a[2] * a[2]))))
    if n == 0:
        raise ModelError(path, "axis must be non-zero")
    return [f32(x / n) for x in a]


def _inertia(v, mass, path):
    """parse_inertia (model.rs:326-400): diag, tensor or solid shape;
    column-major Mat3 (tensor columns (Ixx,Ixy,Ixz), (Ixy,Iyy,Iyz), ...)."""
    kind = _obj(v, path, {"kind", "values", "shape"}).get("kind")
    if kind == "diag":
        d = _vec(v.get("values"), 3, f"{path}.values")
        if not all(x > 0 for x in d):
            raise ModelError(f"{path}.values", "diagonal moments must be > 0")
        return (d[0], 0, 0, 0, d[1], 0, 0, 0, d[2])
    if kind == "tensor":
        t = _vec(v.get("values"), 6, f"{path}.values")
        ixx, iyy, izz, ixy, ixz, iyz = t
        m = np.array([[ixx, ixy, ixz], [ixy, iyy, iyz], [ixz, iyz, izz]], dtype=np.float64)
        if abs(np.linalg.det(m)) == 0:
            raise ModelError(f"{path}.values", "inertia tensor is singular")
        return (ixx, ixy, ixz, ixy, iyy, iyz, ixz, iyz, izz)
    if kind == "solid":
        sh = v.get("shape")
        sk = _obj(sh, f"{path}.shape", {"kind", "half_extents", "radius", "half_height", "semi_axes"}).get("kind")
        if sk == "box":
            h = _vec(sh.get("half_extents"), 3, f"{path}.shape.half_extents")
            hx2, hy2, hz2 = f32(h[0] * h[0]), f32(h[1] * h[1]), f32(h[2] * h[2])
            t = f32(mass / f32(3.0))
            return (f32(t * f32(hy2 + hz2)), 0, 0, 0, f32(t * f32(hx2 + hz2)), 0, 0, 0, f32(t * f32(hx2 + hy2)))
        if sk == "sphere":
            r = _num(sh.get("radius"), f"{path}.shape.radius", positive=True)
            i = f32(f32(f32(f32(2.0) / f32(5.0)) * mass) * r) * r
            return (i, 0, 0, 0, i, 0, 0, 0, i)
        if sk == "capsule":
            r = _num(sh.get("radius"), f"{path}.shape.radius", positive=True)
            hh = _num(sh.get("half_height"), f"{path}.shape.half_height", nonneg=True)
            ixx, izz = _capsule_inertia32(mass, r, hh)
            return (ixx, 0, 0, 0, ixx, 0, 0, 0, izz)
        raise ModelError(f"{path}.shape.kind", f"solid {sk!r} inertia is not built on the GPU path")
    raise ModelError(f"{path}.kind", "expected diag, tensor or solid")


def _capsule_inertia32(m, r, h):
    """solid_capsule_inertia (geom.rs:966-981) in float32."""
    vol_cyl = f32(f32(r * r) * f32(f32(2.0) * h))
    vol_caps = f32(f32(f32(f32(f32(4.0) / f32(3.0)) * r) * r) * r)
    total = f32(vol_cyl + vol_caps)
    m_c = f32(m * f32(vol_cyl / total))
    m_s = f32(m - m_c)
    i_zz = f32(f32(f32(f32(f32(0.5) * m_c) * r) * r) + f32(f32(f32(f32(f32(2.0) / f32(5.0)) * m_s) * r) * r))
    offset = f32(h + f32(f32(f32(3.0) / f32(8.0)) * r))
    hemi = f32(f32(f32(f32(f32(83.0) / f32(320.0)) * m_s) * r) * r)
    t1 = f32(f32(f32(f32(1.0) / f32(12.0)) * m_c) * f32(f32(f32(f32(3.0) * r) * r) + f32(f32(f32(4.0) * h) * h)))
    i_xx = f32(f32(t1 + hemi) + f32(f32(m_s * offset) * offset))
    return i_xx, i_zz


def _solref(v, path):
    if v is None:
        return (f32(0.02), f32(1.0))
    _obj(v, path, {"timeconst", "dampratio"})
    return (_num(v.get("timeconst", 0.02), f"{path}.timeconst"), _num(v.get("dampratio", 1.0), f"{path}.dampratio"))


def _solimp(v, path):
    if v is None:
        return (f32(0.9), f32(0.95), f32(0.001), f32(0.5), 2)
    _obj(v, path, {"dmin", "dmax", "width", "midpoint", "power"})
    power = v.get("power", 2)
    if not isinstance(power, int) or isinstance(power, bool) or power < 1:
        raise ModelError(f"{path}.power", "expected a positive integer")
    return (_num(v.get("dmin", 0.9), f"{path}.dmin"), _num(v.get("dmax", 0.95), f"{path}.dmax"),
            _num(v.get("width", 0.001), f"{path}.width"), _num(v.get("midpoint", 0.5), f"{path}.midpoint"), power)


UNBUILT = {"actuators": "actuators", "equality": "equality constraints",
           "tendons": "tendons", "hfields": "heightfields",
           "contact_pairs": "explicit contact-pair lists", "sensors": "sensors"}


def load(src) -> scenes.Scene:
    """Load a scene from a JSON file path or a JSON text."""
    if isinstance(src, (str, os.PathLike)) and os.path.exists(src):
        name = os.path.splitext(os.path.basename(src))[0]
        with open(src) as f:
            text = f.read()
    else:
        name, text = "json", src
    try:
        root = json.loads(text, object_pairs_hook=_no_duplicates)
    except json.JSONDecodeError as e:
        raise ModelError("<json>", str(e)) from None
    top = _obj(root, "$", {"version", "gravity", "timestep", "integrator", "bodies", "trees", "meshes", "hfields",
                           "geoms", "sites", "actuators", "contact_pairs", "solver", "equality", "tendons",
                           "sensors"})
    if "version" in top and top["version"] != "1":
        raise ModelError("$.version", 'must be "1"')
    for key, what in UNBUILT.items():
        if top.get(key):
            raise ModelError(f"$.{key}", f"{what} are not built on the GPU path")
    gravity = _vec(top.get("gravity", [0.0, 0.0, -9.81]), 3, "$.gravity")
    dt = _num(top.get("timestep", 0.005), "$.timestep", positive=True)
    integ = {"RK4": abi.INTEGRATOR_RK4, "Euler": abi.INTEGRATOR_EULER,
             "implicitfast": abi.INTEGRATOR_IMPLICITFAST}.get(top.get("integrator", "RK4"))
    if integ is None:
        raise ModelError("$.integrator", "expected RK4, Euler or implicitfast")
    solver = _obj(top.get("solver", {}), "$.solver", {"mode", "iterations", "cone"})
    mode = {"penalty": abi.SOLVER_PENALTY, "pgs": abi.SOLVER_PGS, "newton": abi.SOLVER_NEWTON}.get(
        solver.get("mode", "penalty"))
    if mode is None:
        raise ModelError("$.solver.mode", "expected penalty, pgs or newton")
    iterations = solver.get("iterations", 20)
    if not isinstance(iterations, int) or isinstance(iterations, bool) or iterations < 1:
        raise ModelError("$.solver.iterations", "expected a positive integer")
    cone = {"pyramidal": abi.CONE_PYRAMIDAL, "elliptic": abi.CONE_ELLIPTIC}.get(solver.get("cone", "pyramidal"))
    if cone is None:
        raise ModelError("$.solver.cone", "expected pyramidal or elliptic")

    trees = top.get("trees", [])
    bodies_v = top.get("bodies", [])
    if not isinstance(bodies_v, list):
        raise ModelError("$.bodies", "expected an array")
    bodies, body_index = _bodies(bodies_v)
    tree_index, indexes, loaded = {}, [], []
    if trees or not bodies:
        if not isinstance(trees, list) or not trees:
            raise ModelError("$.trees", "expected a non-empty array of trees")
        for k, t in enumerate(trees):
            tp = f"$.trees[{k}]"
            tree = _obj(t, tp, {"name", "self_collide", "links"})
            tname = tree.get("name", "tree")
            if tname in tree_index:
                raise ModelError(f"{tp}.name", "expected a unique name")
            collide = tree.get("self_collide", False)
            if not isinstance(collide, bool):
                raise ModelError(f"{tp}.self_collide", "expected a bool")
            tlinks, tidx = _links(tree.get("links", []), f"{tp}.links")
            tree_index[tname] = k
            indexes.append(tidx)
            loaded.append((tlinks, 1 if collide else 0))
    links, self_collide = (loaded[0] if loaded else ([], 0))
    index = indexes[0] if indexes else {}
    meshes, mesh_index = _meshes(top.get("meshes", []))
    geoms = _geoms(top.get("geoms", []), tree_index, indexes, body_index, mesh_index)
    sites = top.get("sites", [])
    opts = dict(dt=float(dt), gravity=tuple(float(g) for g in gravity), integrator=integ, solver=mode,
                iterations=iterations, cone=cone)
    sc = scenes.Scene(name, opts, links, geoms, self_collide, None, f"loaded from JSON ({name})",
                      bodies=bodies, more_trees=loaded[1:], meshes=meshes)
    sc.extra["sites"] = sites
    sc.extra["link_names"] = list(index)
    sc.extra["body_names"] = list(body_index)
    if bodies and not loaded:
        q0 = np.array([x for b in bodies for x in (*b["position"], *b["orientation"])])
        v0 = np.array([x for b in bodies for x in (*b["linear_velocity"], *b["angular_velocity_body"])])
        sc.sampler = lambda n, seed: (np.tile(q0, (n, 1)), np.tile(v0, (n, 1)))
    return sc


def _bodies(arr):
    """parse_body (model.rs:1143-1186): mass, inertia, pose, velocity."""
    out, index = [], {}
    for i, b in enumerate(arr):
        p = f"$.bodies[{i}]"
        _obj(b, p, {"name", "mass", "inertia", "pose", "velocity", "rolling_friction"})
        name = b.get("name")
        if not isinstance(name, str) or not name or name in index:
            raise ModelError(f"{p}.name", "expected a unique name")
        if "mass" not in b:
            raise ModelError(f"{p}.mass", "required")
        mass = _num(b["mass"], f"{p}.mass", positive=True)
        if "inertia" not in b:
            raise ModelError(f"{p}.inertia", "required")
        inertia = _inertia(b["inertia"], mass, f"{p}.inertia")
        pos, quat = _pose(b.get("pose"), f"{p}.pose")
        if b.get("rolling_friction") is not None:
            raise ModelError(f"{p}.rolling_friction", "rolling friction is not built on the GPU path")
        lin, ang = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
        if "velocity" in b:
            vel = _obj(b["velocity"], f"{p}.velocity", {"linear", "angular_body"})
            if "linear" in vel:
                lin = _vec(vel["linear"], 3, f"{p}.velocity.linear")
            if "angular_body" in vel:
                ang = _vec(vel["angular_body"], 3, f"{p}.velocity.angular_body")
        index[name] = i
        out.append(scenes.body(float(mass), tuple(float(x) for x in inertia), tuple(float(x) for x in pos),
                               tuple(float(x) for x in quat), tuple(float(x) for x in lin),
                               tuple(float(x) for x in ang)))
    return out, index


def _no_duplicates(pairs):
    out = {}
    for k, v in pairs:
        if k in out:
            raise ModelError(k, "duplicate key")
        out[k] = v
    return out


def _links(arr, path):
    if not isinstance(arr, list) or not arr:
        raise ModelError(path, "expected a non-empty array of links")
    out: List[Dict[str, Any]] = []
    index: Dict[str, int] = {}
    for i, l in enumerate(arr):
        p = f"{path}[{i}]"
        _obj(l, p, {"name", "parent", "joint", "joint_offset_in_parent", "joint_offset_in_child", "mass", "inertia"})
        name = l.get("name")
        if not isinstance(name, str) or name in index:
            raise ModelError(f"{p}.name", "expected a unique name")
        parent = l.get("parent")
        if i == 0:
            if parent is not None:
                raise ModelError(f"{p}.parent", "the root has no parent")
            pi = -1
        else:
            if parent not in index:
                raise ModelError(f"{p}.parent", "must name an earlier link")
            pi = index[parent]
        j = _obj(l.get("joint"), f"{p}.joint", {"kind", "axis", "range", "damping", "armature", "limit"})
        kind = j.get("kind")
        kinds = {"free": abi.JOINT_FREE, "fixed": abi.JOINT_FIXED, "hinge": abi.JOINT_HINGE, "slide": abi.JOINT_SLIDE}
        if kind == "ball":
            raise ModelError(f"{p}.joint.kind", "ball joints are not built on the GPU path")
        if kind not in kinds:
            raise ModelError(f"{p}.joint.kind", "expected free, fixed, hinge or slide")
        if i == 0 and kind not in ("free", "fixed"):
            raise ModelError(f"{p}.joint.kind", "the root joint must be free or fixed")
        if i > 0 and kind == "free":
            raise ModelError(f"{p}.joint.kind", "free joints only at the root")
        single = kind in ("hinge", "slide")
        axis = _normalized(_vec(j.get("axis"), 3, f"{p}.joint.axis"), f"{p}.joint.axis") if single else [0, 0, 0]
        rng = None
        if "range" in j:
            if not single:
                raise ModelError(f"{p}.joint.range", "only hinges and slides take a range")
            lo, hi = _vec(j["range"], 2, f"{p}.joint.range")
            if not lo < hi:
                raise ModelError(f"{p}.joint.range", "lo < hi required")
            rng = (float(lo), float(hi))
        damping = _num(j.get("damping", 0.0), f"{p}.joint.damping", nonneg=True)
        armature = _num(j.get("armature", 0.0), f"{p}.joint.armature", nonneg=True)
        lsr = lsi = lk = lc = None
        if "limit" in j:
            # model.rs:679-735: stiffness and damping required, solref/solimp optional
            lim = _obj(j["limit"], f"{p}.joint.limit", {"stiffness", "damping", "solref", "solimp"})
            if "stiffness" not in lim or "damping" not in lim:
                raise ModelError(f"{p}.joint.limit", "stiffness and damping are required")
            lk = float(_num(lim["stiffness"], f"{p}.joint.limit.stiffness", nonneg=True))
            lc = float(_num(lim["damping"], f"{p}.joint.limit.damping", nonneg=True))
            if "solref" in lim:
                lsr = tuple(float(x) for x in _solref(lim["solref"], f"{p}.joint.limit.solref"))
            if "solimp" in lim:
                si = _solimp(lim["solimp"], f"{p}.joint.limit.solimp")
                lsi = tuple(float(x) for x in si[:4]) + (si[4],)
        mass = _num(l.get("mass"), f"{p}.mass", positive=True)
        inertia = _inertia(l.get("inertia"), mass, f"{p}.inertia")
        # The root's parent offset is its world pose; other offsets carry no
        # rotation (docs/model-format.md, "links").
        ppos, pquat = _pose(l.get("joint_offset_in_parent"), f"{p}.joint_offset_in_parent", allow_orientation=i == 0)
        cpos, cquat = _pose(l.get("joint_offset_in_child"), f"{p}.joint_offset_in_child", allow_orientation=False)
        d = scenes.link(pi, kinds[kind], axis=tuple(float(x) for x in axis),
                        parent_pos=tuple(float(x) for x in ppos), parent_quat=tuple(float(x) for x in pquat),
                        child_pos=tuple(float(x) for x in cpos), mass=float(mass),
                        inertia=tuple(float(x) for x in inertia),
                        damping=float(damping) if kind != "free" else 0.0, armature=float(armature),
                        free_damping=float(damping) if kind == "free" else 0.0, range=rng, limit_solref=lsr,
                        limit_solimp=lsi, limit_stiffness=lk, limit_damping=lc)
        out.append(d)
        index[name] = i
    return out, index


def _meshes(arr):
    """parse_mesh_asset (model.rs:1640-1699): named convex meshes, checked as
    ConvexMesh::validate does (geom.rs:318-340)."""
    if not isinstance(arr, list):
        raise ModelError("$.meshes", "expected an array")
    out, index = [], {}
    for i, m in enumerate(arr):
        p = f"$.meshes[{i}]"
        _obj(m, p, {"name", "vertices", "faces"})
        name = m.get("name")
        if not isinstance(name, str) or not name:
            raise ModelError(f"{p}.name", "mesh name must not be empty")
        if name in index:
            raise ModelError(p, f'duplicate mesh name "{name}"')
        verts = m.get("vertices")
        if not isinstance(verts, list):
            raise ModelError(f"{p}.vertices", "expected an array")
        v = [_vec(x, 3, f"{p}.vertices[{k}]") for k, x in enumerate(verts)]
        faces = m.get("faces")
        if not isinstance(faces, list):
            raise ModelError(f"{p}.faces", "expected an array")
        f = []
        for k, t in enumerate(faces):
            if not isinstance(t, list) or len(t) != 3:
                raise ModelError(f"{p}.faces[{k}]", "face must have exactly 3 vertex indices")
            tri = []
            for c, e in enumerate(t):
                n = _num(e, f"{p}.faces[{k}][{c}]")
                if n < 0 or n != math.floor(n):
                    raise ModelError(f"{p}.faces[{k}][{c}]", "face vertex index must be a non-negative integer")
                tri.append(int(n))
            f.append(tri)
        if len(v) < 4:
            raise ModelError(p, "convex mesh needs >= 4 vertices")
        if len(f) < 4:
            raise ModelError(p, "convex mesh needs >= 4 triangular faces")
        if any(c >= len(v) for tri in f for c in tri):
            raise ModelError(p, "mesh face index out of range")
        index[name] = i
        out.append((np.array(v, dtype=np.float64), np.array(f, dtype=np.int32)))
    return out, index


def _geoms(arr, tree_index, indexes, body_index=None, mesh_index=None):
    body_index = body_index or {}
    mesh_index = mesh_index or {}
    out = []
    shapes = {"plane": abi.SHAPE_PLANE, "sphere": abi.SHAPE_SPHERE, "box": abi.SHAPE_BOX, "capsule": abi.SHAPE_CAPSULE,
              "mesh": abi.SHAPE_MESH}
    for i, g in enumerate(arr):
        p = f"$.geoms[{i}]"
        _obj(g, p, {"name", "shape", "attach", "local_offset", "local_orientation", "friction", "solref", "solimp",
                    "condim", "margin", "gap", "torsional_friction", "rolling_friction"})
        sh = _obj(g.get("shape"), f"{p}.shape", {"kind", "radius", "half_extents", "half_height", "semi_axes",
                                                  "mesh", "hfield"})
        kind = sh.get("kind")
        if kind not in shapes:
            raise ModelError(f"{p}.shape.kind", f"{kind!r} geoms are not built on the GPU path")
        if kind == "sphere":
            size = (float(_num(sh.get("radius"), f"{p}.shape.radius", positive=True)), 0.0, 0.0)
        elif kind == "box":
            size = tuple(float(x) for x in _vec(sh.get("half_extents"), 3, f"{p}.shape.half_extents"))
        elif kind == "capsule":
            size = (float(_num(sh.get("radius"), f"{p}.shape.radius", positive=True)),
                    float(_num(sh.get("half_height"), f"{p}.shape.half_height", nonneg=True)), 0.0)
        else:
            size = (0.0, 0.0, 0.0)
        mesh = -1
        if kind == "mesh":
            _obj(sh, f"{p}.shape", {"kind", "mesh"})
            if sh.get("mesh") not in mesh_index:
                raise ModelError(f"{p}.shape.mesh", "unknown mesh name; declare it in the top-level meshes array")
            mesh = mesh_index[sh["mesh"]]
        att = _obj(g.get("attach", {"kind": "static"}), f"{p}.attach", {"kind", "tree", "link", "body"})
        tree, link, body = -1, -1, -1
        if att.get("kind") == "static":
            if kind != "plane":
                raise ModelError(f"{p}.attach", "only plane geoms may be static")
        elif att.get("kind") == "link":
            t = tree_index.get(att.get("tree")) if isinstance(att.get("tree"), str) else None
            if t is None or att.get("link") not in indexes[t]:
                raise ModelError(f"{p}.attach", "must name a tree and one of its links")
            tree, link = t, indexes[t][att["link"]]
        elif att.get("kind") == "body":
            if att.get("body") not in body_index:
                raise ModelError(f"{p}.attach.body", "must name one of the bodies")
            body = body_index[att["body"]]
        else:
            raise ModelError(f"{p}.attach.kind", "expected static, link or body")
        pos = tuple(float(x) for x in _vec(g.get("local_offset", [0, 0, 0]), 3, f"{p}.local_offset"))
        quat = tuple(float(x) for x in _quat_normalized(_vec(g.get("local_orientation", [0, 0, 0, 1]), 4,
                                                             f"{p}.local_orientation")))
        condim = g.get("condim", 3)
        if condim not in (1, 3, 4, 6):
            raise ModelError(f"{p}.condim", "expected 1, 3, 4 or 6")
        si = _solimp(g.get("solimp"), f"{p}.solimp")
        out.append(scenes.geom(shapes[kind], tree=tree, link=link, body=body, size=size, pos=pos, quat=quat,
                               friction=float(_num(g.get("friction", 0.5), f"{p}.friction", nonneg=True)),
                               margin=float(_num(g.get("margin", 0.0), f"{p}.margin", nonneg=True)),
                               gap=float(_num(g.get("gap", 0.0), f"{p}.gap", nonneg=True)),
                               solref=tuple(float(x) for x in _solref(g.get("solref"), f"{p}.solref")),
                               solimp=tuple(float(x) for x in si[:4]), power=si[4], condim=condim, mesh=mesh))
    return out
