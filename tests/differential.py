"""Helpers of the MuJoCo differential checks (newt/tests/differential.rs):
newt's MJCF scenarios with MuJoCo 3.11's recorded trajectories
(tests/golden/mujoco/, copies of newt/tests/references/ scenario files,
fixtures and scenarios.json), the state remap between MuJoCo's qpos/qvel and
this library's, and newt's per-scenario tolerance windows."""
import json
import os
import struct

import numpy as np

from tooling_amd import abi, mjcf

HERE = os.path.dirname(os.path.abspath(__file__))
MJ = os.path.join(HERE, "golden", "mujoco")

# (scenario, fixture suffix, integrator override, solver override, qpos tol,
# qvel tol): the bounds newt's differential tests assert (differential.rs:
# 66-231 tolerance(), 1272-1293 matched_integrator_tolerance(), 1411-1444
# the tree-chain rows, 1619-1650 the Newton rows).
CASES = [
    ("ballistic", "", None, None, 3.0e-5, 8.0e-5),
    ("ballistic", "_euler", abi.INTEGRATOR_EULER, None, 3.0e-5, 8.0e-5),
    ("tumble", "", None, None, 4.0e-3, 5.0e-2),
    ("double_pendulum", "", None, None, 5.0e-7, 2.0e-6),
    ("double_pendulum", "_euler", abi.INTEGRATOR_EULER, None, 4.0e-7, 1.0e-6),
    ("floating_base", "", None, None, 6.0e-6, 6.0e-6),
    ("sphere_drop", "", None, None, 1.0e-2, 1.0),
    ("sphere_drop", "_euler", abi.INTEGRATOR_EULER, None, 1.5e-2, 9.0e-1),
    ("sphere_drop", "_newton_euler", abi.INTEGRATOR_EULER, abi.SOLVER_NEWTON, 2.0e-2, 1.2),
    ("sphere_drop_stiff", "", None, None, 8.0e-2, 8.0e-1),
    ("sphere_drop_soft", "", None, None, 7.0e-2, 8.0e-1),
    ("box_stack", "", None, None, 5.0e-2, 5.0e-1),
    ("box_stack", "_euler", abi.INTEGRATOR_EULER, None, 2.0e-2, 5.0e-1),
    ("box_stack", "_newton_euler", abi.INTEGRATOR_EULER, abi.SOLVER_NEWTON, 5.0e-2, 1.0),
    ("joint_limit_swing", "", None, None, 1.5e-1, 1.2),
    ("tree_chain_contact", "_pgs_euler", abi.INTEGRATOR_EULER, abi.SOLVER_PGS, 7.0e-2, 0.7),
    ("tree_chain_contact", "_newton_euler", abi.INTEGRATOR_EULER, abi.SOLVER_NEWTON, 8.0e-2, 1.3),
]


def case_id(c):
    return c[0] + c[1]


def spec(name):
    with open(os.path.join(MJ, "scenarios.json")) as f:
        return next(s for s in json.load(f)["scenarios"] if s["name"] == name)


def read_fixture(path):
    """NEWTDIF1 (differential.rs:239-302): provenance, then per sample the
    step, qpos and qvel in MuJoCo's layout."""
    b = open(path, "rb").read()
    assert b[:8] == b"NEWTDIF1"
    o = 8
    (plen,) = struct.unpack_from("<I", b, o)
    o += 4 + plen
    nq, nv, stride, ns, nsteps = struct.unpack_from("<5I", b, o)
    o += 20
    qs, vs = [], []
    for _ in range(ns):
        o += 4
        qs.append(struct.unpack_from(f"<{nq}d", b, o))
        o += 8 * nq
        vs.append(struct.unpack_from(f"<{nv}d", b, o))
        o += 8 * nv
    assert o == len(b)
    return stride, nsteps, np.array(qs), np.array(vs)


def scene(c):
    """The scenario's Scene through the MJCF subset loader, with the case's
    integrator and solver overrides."""
    name, _suffix, integ, solver = c[:4]
    sc = mjcf.load(os.path.join(MJ, spec(name)["mjcf"]))
    if integ is not None:
        sc.options["integrator"] = integ
    if solver is not None:
        sc.options["solver"] = solver
    return sc


def _rot(q, v):
    x, y, z, w = q
    u = np.array([x, y, z])
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def _rot_inv(q, v):
    return _rot((-q[0], -q[1], -q[2], q[3]), v)


def _joints(sc):
    return [l["joint"] for links in [sc.links] + [t for t, _ in sc.more_trees] for l in links]


def from_mujoco(sc, q_mj, v_mj):
    """World::apply_mujoco_qpos / apply_mujoco_qvel (world.rs:519-620): free
    bodies (position, wxyz) -> (position, xyzw), velocities as they are; a
    free root's velocity (v_world, w_body) -> (w_body, v_body)."""
    q, v = [], []
    qi = vi = 0
    for j in _joints(sc):
        if j == abi.JOINT_FREE:
            p = q_mj[qi:qi + 3]
            w, x, y, z = q_mj[qi + 3:qi + 7]
            quat = np.array([x, y, z, w]) / np.linalg.norm([x, y, z, w])
            q += [*p, *quat]
            vb = _rot_inv(quat, np.array(v_mj[vi:vi + 3]))
            v += [*v_mj[vi + 3:vi + 6], *vb]
            qi, vi = qi + 7, vi + 6
        elif j != abi.JOINT_FIXED:
            q.append(q_mj[qi])
            v.append(v_mj[vi])
            qi, vi = qi + 1, vi + 1
    for _b in sc.bodies:
        p = q_mj[qi:qi + 3]
        w, x, y, z = q_mj[qi + 3:qi + 7]
        quat = np.array([x, y, z, w]) / np.linalg.norm([x, y, z, w])
        q += [*p, *quat]
        v += list(v_mj[vi:vi + 6])
        qi, vi = qi + 7, vi + 6
    return np.array(q), np.array(v)


def to_mujoco(sc, q, v):
    """extract_qpos / extract_qvel (differential.rs:548-640)."""
    qm, vm = [], []
    qi = vi = 0
    for j in _joints(sc):
        if j == abi.JOINT_FREE:
            quat = q[qi + 3:qi + 7]
            qm += [*q[qi:qi + 3], quat[3], quat[0], quat[1], quat[2]]
            vm += [*_rot(quat, np.array(v[vi + 3:vi + 6])), *v[vi:vi + 3]]
            qi, vi = qi + 7, vi + 6
        elif j != abi.JOINT_FIXED:
            qm.append(q[qi])
            vm.append(v[vi])
            qi, vi = qi + 1, vi + 1
    for _b in sc.bodies:
        quat = q[qi + 3:qi + 7]
        qm += [*q[qi:qi + 3], quat[3], quat[0], quat[1], quat[2]]
        vm += list(v[vi:vi + 6])
        qi, vi = qi + 7, vi + 6
    return np.array(qm), np.array(vm)


def initial_state(c, sc):
    """The scenario's initial state in this library's layout: the model's
    own (bodies and free roots at their MJCF poses, at rest) with
    scenarios.json's init_qpos / init_qvel applied as newt applies them."""
    s = spec(c[0])
    q0, v0 = sc.initial_state(1, 0)
    q_mj, v_mj = to_mujoco(sc, q0[0], v0[0])
    if s.get("init_qpos") is not None:
        q_mj = np.array(s["init_qpos"], dtype=np.float64)
    if s.get("init_qvel") is not None:
        v_mj = np.array(s["init_qvel"], dtype=np.float64)
    return from_mujoco(sc, q_mj, v_mj)


def divergence(sc, fix_q, fix_v, samples):
    """compare_and_measure (differential.rs:654-756): the largest |newt -
    MuJoCo| over samples, free-body quaternions sign-aligned first."""
    nb = len(sc.bodies)
    n_tree_q = len(fix_q[0]) - 7 * nb
    dq = dv = 0.0
    for k, (q, v) in enumerate(samples):
        q = q.copy()
        for b in range(nb):
            s = n_tree_q + 7 * b + 3
            if np.dot(q[s:s + 4], fix_q[k][s:s + 4]) < 0:
                q[s:s + 4] = -q[s:s + 4]
        dq = max(dq, float(np.abs(q - fix_q[k]).max()))
        dv = max(dv, float(np.abs(v - fix_v[k]).max()))
    return dq, dv
