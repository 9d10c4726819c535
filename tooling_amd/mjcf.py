"""The MJCF subset newt loads (newt/src/mjcf.rs) for the models the GPU step
builds.

`load(path_or_text)` returns a :class:`tooling_amd.scenes.Scene` that
:class:`tooling_amd.newt.Model` compiles, following newt's mapping:
- `<option>` timestep, gravity, integrator (RK4 / Euler / implicitfast),
  solver (PGS / Newton), cone, iterations (mjcf.rs:710-790); absent values
  keep World::new's (dt 5 ms, RK4, penalty contacts, 20 iterations);
- `<compiler angle>`; top-level `<default>` geom attributes (friction,
  solref, solimp, condim, margin, gap, contype, conaffinity);
- a top-level body with a `<freejoint>` and no nested bodies becomes a free
  Body (World::bodies, mjcf.rs:889-1027); any other top-level body a tree
  whose root is Fixed (no joint) or Free (`<freejoint>`) and whose links are
  the nested bodies with one hinge or slide joint each or none (Fixed)
  (mjcf.rs:1029-1330): joint_offset_in_parent = body pos + joint pos,
  joint_offset_in_child = joint pos; trees collide with nothing of their
  own (World::disabled_self_collision);
- `<inertial>` mass with `diaginertia` or `fullinertia` at the body origin;
- geoms plane, sphere, box and capsule (size or fromto), with pos / quat,
  friction (the sliding coefficient), solref, solimp, condim, margin, gap,
  contype / conaffinity as collision group / mask (mjcf.rs:1891-2030);
  friction defaults to 0.5 as in newt.
Numbers are rounded to float32 as newt parses them.  Anything outside the
subset raises `MjcfError` naming the element.  The free root's q holds the
body pose, and a body's `pos`/`quat` give its initial state.
"""
from __future__ import annotations

import math
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import abi, scenes
from .model_json import _quat_normalized, f32


class MjcfError(ValueError):
    def __init__(self, path: str, message: str):
        super().__init__(f"{path}: {message}")
        self.path = path


def _floats(s, path, n=None):
    try:
        out = [f32(float(x)) for x in s.split()]
    except ValueError:
        raise MjcfError(path, f"expected numbers, got {s!r}") from None
    if n is not None and len(out) != n:
        raise MjcfError(path, f"expected {n} numbers, got {len(out)}")
    return out


def _quat_wxyz(s, path):
    """parse_quat_wxyz (mjcf.rs:296-307): (x, y, z, w), renormalized."""
    w, x, y, z = _floats(s, path, 4)
    if w * w + x * x + y * y + z * z < 1e-12:
        raise MjcfError(path, "quaternion has zero norm")
    return tuple(float(v) for v in _quat_normalized([x, y, z, w]))


class _Loader:
    def __init__(self, root):
        self.angle = 1.0
        self.opts = dict(dt=0.005, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_RK4,
                         solver=abi.SOLVER_PENALTY, iterations=20, cone=abi.CONE_PYRAMIDAL)
        self.geom_default = {}
        self.trees = []        # (links, geoms relative to the tree) per tree
        self.bodies = []
        self.geoms = []
        self.q_trees = []
        for child in root:
            tag = child.tag
            if tag == "compiler":
                ang = child.get("angle", "degree")
                self.angle = 1.0 if ang == "radian" else math.pi / 180.0
            elif tag == "option":
                self.option(child)
            elif tag == "default":
                for d in child:
                    if d.tag != "geom":
                        raise MjcfError("default", f"<{d.tag}> defaults are not in the subset")
                    self.geom_default.update(d.attrib)
            elif tag == "worldbody":
                self.worldbody(child)
            elif tag in ("asset", "size", "visual", "statistic"):
                if tag == "asset" and len(child):
                    raise MjcfError("asset", "assets are not in the subset")
            else:
                raise MjcfError(tag, f"<{tag}> is not in the subset")

    def option(self, e):
        for k, v in e.attrib.items():
            if k == "timestep":
                self.opts["dt"] = float(_floats(v, "option.timestep", 1)[0])
            elif k == "gravity":
                self.opts["gravity"] = tuple(float(x) for x in _floats(v, "option.gravity", 3))
            elif k == "integrator":
                m = {"RK4": abi.INTEGRATOR_RK4, "rk4": abi.INTEGRATOR_RK4, "Euler": abi.INTEGRATOR_EULER,
                     "euler": abi.INTEGRATOR_EULER, "implicitfast": abi.INTEGRATOR_IMPLICITFAST}
                if v not in m:
                    raise MjcfError("option.integrator", f"{v!r} is not supported")
                self.opts["integrator"] = m[v]
            elif k == "solver":
                m = {"PGS": abi.SOLVER_PGS, "pgs": abi.SOLVER_PGS, "Newton": abi.SOLVER_NEWTON,
                     "newton": abi.SOLVER_NEWTON}
                if v not in m:
                    raise MjcfError("option.solver", f"{v!r} is not supported (PGS or Newton)")
                self.opts["solver"] = m[v]
            elif k == "cone":
                m = {"pyramidal": abi.CONE_PYRAMIDAL, "elliptic": abi.CONE_ELLIPTIC}
                if v not in m:
                    raise MjcfError("option.cone", f"{v!r} is not supported")
                self.opts["cone"] = m[v]
            elif k == "iterations":
                n = int(v)
                if n < 1:
                    raise MjcfError("option.iterations", "must be >= 1")
                self.opts["iterations"] = n
            else:
                raise MjcfError("option", f"attribute {k!r} is not in the subset")

    # -- geoms ------------------------------------------------------------
    def geom(self, e, path, tree=-1, link=-1, body=-1):
        a = dict(self.geom_default)
        a.update(e.attrib)
        ty = a.get("type", "sphere")
        size = _floats(a["size"], f"{path}.size") if "size" in a else []
        pos = tuple(float(x) for x in _floats(a.get("pos", "0 0 0"), f"{path}.pos", 3))
        quat = _quat_wxyz(a["quat"], f"{path}.quat") if "quat" in a else scenes.IDENTITY_Q
        if ty == "plane":
            shape, sz = abi.SHAPE_PLANE, (0.0, 0.0, 0.0)
        elif ty == "sphere":
            if not size:
                raise MjcfError(path, 'sphere requires size="radius"')
            shape, sz = abi.SHAPE_SPHERE, (float(size[0]), 0.0, 0.0)
        elif ty == "box":
            if len(size) != 3:
                raise MjcfError(path, "box size must have 3 numbers")
            shape, sz = abi.SHAPE_BOX, tuple(float(x) for x in size)
        elif ty == "capsule":
            if "fromto" in a:
                ft = _floats(a["fromto"], f"{path}.fromto", 6)
                p0, p1 = np.array(ft[:3], dtype=np.float32), np.array(ft[3:], dtype=np.float32)
                d = p1 - p0
                half = f32(np.sqrt(f32(np.dot(d, d))) * f32(0.5))
                pos = tuple(float(x) for x in (p0 + p1) * f32(0.5))
                axis = d / f32(2.0 * half) if half > 0 else np.array([0, 0, 1], dtype=np.float32)
                zc = np.cross([0.0, 0.0, 1.0], axis)
                c = float(axis[2])
                if np.linalg.norm(zc) < 1e-12:
                    quat = scenes.IDENTITY_Q if c > 0 else (1.0, 0.0, 0.0, 0.0)
                else:
                    quat = scenes.quat_axis_angle(tuple(zc), math.acos(max(-1.0, min(1.0, c))))
                shape, sz = abi.SHAPE_CAPSULE, (float(size[0]), float(half), 0.0)
            else:
                if len(size) < 2:
                    raise MjcfError(path, 'capsule requires size="radius half_height" or fromto')
                shape, sz = abi.SHAPE_CAPSULE, (float(size[0]), float(size[1]), 0.0)
        else:
            raise MjcfError(path, f"geom type {ty!r} is not built on the GPU path")
        fr = _floats(a["friction"], f"{path}.friction") if "friction" in a else [f32(0.5)]
        if len(fr) > 1 and any(x != 0 for x in fr[1:]):
            raise MjcfError(path, "torsional and rolling friction are not built")
        solref = tuple(float(x) for x in _floats(a["solref"], f"{path}.solref", 2)) if "solref" in a else (0.02, 1.0)
        if "solimp" in a:
            si = _floats(a["solimp"], f"{path}.solimp")
            if len(si) not in (3, 5):
                raise MjcfError(path, "solimp must have 3 or 5 numbers")
            solimp = (float(si[0]), float(si[1]), float(si[2]), float(si[3]) if len(si) > 3 else 0.5)
            power = int(si[4]) if len(si) > 4 else 2
        else:
            solimp, power = (0.9, 0.95, 0.001, 0.5), 2
        condim = int(a.get("condim", 3))
        if tree < 0 and body < 0 and shape != abi.SHAPE_PLANE:
            pass  # static primitives are allowed (World::add_geom with no attachment)
        g = scenes.geom(shape, tree=tree, link=link, body=body, size=sz, pos=pos, quat=quat,
                        friction=float(fr[0]), margin=float(_floats(a.get("margin", "0"), f"{path}.margin", 1)[0]),
                        gap=float(_floats(a.get("gap", "0"), f"{path}.gap", 1)[0]), solref=solref, solimp=solimp,
                        power=power, condim=condim, group=int(a.get("contype", 1)), mask=int(a.get("conaffinity", 1)))
        self.geoms.append(g)

    # -- bodies -----------------------------------------------------------
    def pose(self, e, path):
        pos = tuple(float(x) for x in _floats(e.get("pos", "0 0 0"), f"{path}.pos", 3))
        if "quat" in e.attrib:
            q = _quat_wxyz(e.get("quat"), f"{path}.quat")
        elif "axisangle" in e.attrib:
            aa = _floats(e.get("axisangle"), f"{path}.axisangle", 4)
            q = scenes.quat_axis_angle(tuple(float(x) for x in aa[:3]), float(aa[3]) * self.angle)
        elif "euler" in e.attrib:
            raise MjcfError(path, "euler orientations are not in the subset")
        else:
            q = scenes.IDENTITY_Q
        return pos, q

    @staticmethod
    def inertial(e, path):
        ine = [c for c in e if c.tag == "inertial"]
        if len(ine) != 1:
            raise MjcfError(path, "a body needs exactly one <inertial>")
        i = ine[0]
        if any(abs(x) > 0 for x in _floats(i.get("pos", "0 0 0"), f"{path}.inertial.pos", 3)):
            raise MjcfError(f"{path}.inertial", "inertial pos must be [0 0 0] (COM at the body origin)")
        if "quat" in i.attrib:
            raise MjcfError(f"{path}.inertial", "inertial quat is not supported")
        mass = float(_floats(i.get("mass"), f"{path}.inertial.mass", 1)[0])
        if "diaginertia" in i.attrib:
            d = [float(x) for x in _floats(i.get("diaginertia"), f"{path}.inertial.diaginertia", 3)]
            inertia = (d[0], 0.0, 0.0, 0.0, d[1], 0.0, 0.0, 0.0, d[2])
        elif "fullinertia" in i.attrib:
            xx, yy, zz, xy, xz, yz = (float(x) for x in _floats(i.get("fullinertia"), f"{path}.fullinertia", 6))
            inertia = (xx, xy, xz, xy, yy, yz, xz, yz, zz)
        else:
            raise MjcfError(f"{path}.inertial", "diaginertia or fullinertia is required")
        return mass, inertia

    def worldbody(self, wb):
        for c in wb:
            path = f"worldbody.{c.tag}[{c.get('name', '')}]"
            if c.tag == "geom":
                self.geom(c, path)
            elif c.tag == "body":
                self.top_body(c, path)
            elif c.tag in ("site", "light", "camera"):
                continue
            else:
                raise MjcfError(path, f"<{c.tag}> is not in the subset")

    def top_body(self, e, path):
        free = [c for c in e if c.tag == "freejoint" or (c.tag == "joint" and c.get("type") == "free")]
        joints = [c for c in e if c.tag in ("joint", "freejoint")]
        children = [c for c in e if c.tag == "body"]
        pos, q = self.pose(e, path)
        mass, inertia = self.inertial(e, path)
        if free and not children:
            if len(joints) != 1:
                raise MjcfError(path, "a free-jointed top-level body must have exactly one free joint")
            b = len(self.bodies)
            self.bodies.append(scenes.body(mass, inertia, pos, q))
            for c in e:
                if c.tag == "geom":
                    self.geom(c, f"{path}.geom", body=b)
            return
        if joints and not free:
            raise MjcfError(path, "a top-level body with a hinge or slide joint is not in the subset")
        t = len(self.trees)
        links = [scenes.link(-1, abi.JOINT_FREE if free else abi.JOINT_FIXED, parent_pos=pos, parent_quat=q,
                             mass=mass, inertia=inertia)]
        self.trees.append(links)
        self.q_trees.append([*pos, *q] if free else [])
        for c in e:
            if c.tag == "geom":
                self.geom(c, f"{path}.geom", tree=t, link=0)
        for c in children:
            self.child_body(c, f"{path}.body[{c.get('name', '')}]", t, 0)

    def child_body(self, e, path, t, parent):
        pos, q = self.pose(e, path)
        if q != scenes.IDENTITY_Q:
            raise MjcfError(path, "non-root body orientation is not supported (identity only)")
        mass, inertia = self.inertial(e, path)
        joints = [c for c in e if c.tag in ("joint", "freejoint")]
        if len(joints) > 1:
            raise MjcfError(path, "one joint per body")
        kind, jpos, extra = abi.JOINT_FIXED, (0.0, 0.0, 0.0), {}
        if joints:
            j = joints[0]
            ty = j.get("type", "hinge")
            if j.tag == "freejoint" or ty == "free":
                raise MjcfError(path, "a free joint is only allowed on a top-level body")
            if ty not in ("hinge", "slide"):
                raise MjcfError(path, f"joint type {ty!r} is not built on the GPU path")
            kind = abi.JOINT_HINGE if ty == "hinge" else abi.JOINT_SLIDE
            jpos = tuple(float(x) for x in _floats(j.get("pos", "0 0 0"), f"{path}.joint.pos", 3))
            extra["axis"] = tuple(float(x) for x in _floats(j.get("axis", "0 0 1"), f"{path}.joint.axis", 3))
            extra["damping"] = float(_floats(j.get("damping", "0"), f"{path}.joint.damping", 1)[0])
            extra["armature"] = float(_floats(j.get("armature", "0"), f"{path}.joint.armature", 1)[0])
            if "range" in j.attrib and j.get("limited", "true") != "false":
                r = _floats(j.get("range"), f"{path}.joint.range", 2)
                scale = self.angle if kind == abi.JOINT_HINGE else 1.0
                extra["range"] = (float(r[0]) * scale, float(r[1]) * scale)
        pp = tuple(float(f32(a) + f32(b)) for a, b in zip(pos, jpos))
        links = self.trees[t]
        me = len(links)
        links.append(scenes.link(parent, kind, parent_pos=pp, child_pos=jpos, mass=mass, inertia=inertia, **extra))
        if kind != abi.JOINT_FIXED:
            self.q_trees[t].append(0.0)
        for c in e:
            if c.tag == "geom":
                self.geom(c, f"{path}.geom", tree=t, link=me)
            elif c.tag == "body":
                self.child_body(c, f"{path}.body[{c.get('name', '')}]", t, me)
            elif c.tag not in ("inertial", "joint", "site"):
                raise MjcfError(path, f"<{c.tag}> is not in the subset")


def load(src) -> scenes.Scene:
    """Load a Scene from an MJCF file path or text."""
    if isinstance(src, (str, os.PathLike)) and os.path.exists(src):
        name = os.path.splitext(os.path.basename(src))[0]
        root = ET.parse(src).getroot()
    else:
        name, root = "mjcf", ET.fromstring(src)
    if root.tag != "mujoco":
        raise MjcfError("<root>", "expected <mujoco>")
    ld = _Loader(root)
    opts = dict(ld.opts)
    trees = ld.trees
    q_tree = [x for qs in ld.q_trees for x in qs]
    q_body = [x for b in ld.bodies for x in (*b["position"], *b["orientation"])]
    v_body = [x for b in ld.bodies for x in (*b["linear_velocity"], *b["angular_velocity_body"])]
    q0 = np.array(q_tree + q_body, dtype=np.float64)
    nv = sum(6 if l["joint"] == abi.JOINT_FREE else 0 if l["joint"] == abi.JOINT_FIXED else 1
             for links in trees for l in links)
    v0 = np.array([0.0] * nv + v_body, dtype=np.float64)
    sc = scenes.Scene(name, opts, trees[0] if trees else [], ld.geoms, 0,
                      lambda n, seed: (np.tile(q0, (n, 1)), np.tile(v0, (n, 1))), f"loaded from MJCF ({name})",
                      bodies=ld.bodies, more_trees=[(links, 0) for links in trees[1:]])
    return sc
