"""ctypes wrapper of oracle/_build/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Never imported by tooling_amd.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    """Compile the oracle with its Makefile (IEEE-strict g++)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        P, I, PD, PI = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)
        sig = {
            "oracle_create": (P, [P, I]),
            "oracle_destroy": (None, [P]),
            "oracle_add_tree": (None, [P, P, I, I]),
            "oracle_add_geom": (None, [P, P]),
            "oracle_add_body": (None, [P, P]),
            "oracle_add_mesh": (I, [P, PD, I, PI, I]),
            "oracle_narrow": (I, [I, P, PD, P, PD, PD, I, PI, I, I, PD]),
            "oracle_finalize": (None, [P, I]),
            "oracle_nq": (I, [P]),
            "oracle_nv": (I, [P]),
            "oracle_set_state": (None, [P, PD, PD]),
            "oracle_get_state": (None, [P, PD, PD]),
            "oracle_get_qacc": (None, [P, PD]),
            "oracle_set_qfrc": (None, [P, PD]),
            "oracle_get_qfrc": (None, [P, PD]),
            "oracle_step": (None, [P, I, I]),
            "oracle_contacts": (I, [P, I, I, PI, PD]),
            "oracle_pairs": (I, [P, I, I, PI]),
            "oracle_detect": (I, [P, I]),
            "oracle_mass_matrix": (None, [P, I, PD]),
            "oracle_dynamics": (None, [P, I, PD, PD, PD]),
            "oracle_golden_bytes": (I, [C.c_char_p, C.POINTER(C.c_uint8), I]),
            "oracle_flops": (None, [C.POINTER(C.c_ulonglong), I]),
            "oracle_broadphase": (I, [P, I, PD, PD]),
            "oracle_reinserts": (C.c_longlong, [P, I, I]),
            "oracle_box_pairs_f32": (I, [I, C.POINTER(C.c_float), I, PI]),
            "oracle_nsin64": (C.c_double, [C.c_double]),
            "oracle_ncos64": (C.c_double, [C.c_double]),
            "oracle_nsin32": (C.c_float, [C.c_float]),
            "oracle_ncos32": (C.c_float, [C.c_float]),
            "oracle_add_sensor": (None, [P, P]),
            "oracle_set_magnetic": (None, [P, PD]),
            "oracle_sensor_dim": (I, [P]),
            "oracle_get_sensordata": (None, [P, PD]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def golden_bytes(name: str) -> bytes:
    buf = (C.c_uint8 * 65536)()
    n = lib().oracle_golden_bytes(name.encode(), buf, len(buf))
    if n < 0:
        raise KeyError(name)
    return bytes(buf[:n])


def box_pairs_f32(boxes):
    """The oracle's f32 dynamic-tree pair set of standalone boxes
    [n, 6] = (min xyz, max xyz): fattened at insertion, sorted (a, b)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n = boxes.shape[0]
    cap = n * (n - 1) // 2 + 1
    out = np.zeros((cap, 2), dtype=np.int32)
    k = lib().oracle_box_pairs_f32(n, boxes.ctypes.data_as(C.POINTER(C.c_float)), cap,
                                   out.ctypes.data_as(C.POINTER(C.c_int)))
    return out[:k].copy()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def narrow_phase(geom_a, pose_a, geom_b, pose_b, mesh=None, precision=1):
    """One narrow-phase call on two geoms (abi.GeomDesc) at world poses
    (x, y, z, qx, qy, qz, qw); mesh geoms use `mesh` = (vertices, faces).
    Returns [n, 7] = (position, normal from B into A, penetration)."""
    verts, faces = mesh if mesh is not None else (np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int32))
    v = np.ascontiguousarray(verts, dtype=np.float64)
    f = np.ascontiguousarray(faces, dtype=np.int32)
    pa = np.ascontiguousarray(pose_a, dtype=np.float64)
    pb = np.ascontiguousarray(pose_b, dtype=np.float64)
    out = np.zeros((8, 7))
    n = lib().oracle_narrow(precision, C.cast(C.pointer(geom_a), C.c_void_p), _dp(pa),
                            C.cast(C.pointer(geom_b), C.c_void_p), _dp(pb), _dp(v), len(v),
                            f.ctypes.data_as(C.POINTER(C.c_int)), len(f), 8, _dp(out))
    return out[:n].copy()


class OracleWorlds:
    """N oracle worlds built from a tooling_amd.scenes.Scene.
    precision: 0 = double (FP64 oracle), 1 = float (newt's arithmetic),
    2 = double with operation counts (see flops())."""

    def __init__(self, scene, n_worlds, precision=0):
        L = lib()
        self.L = L
        self.n = n_worlds
        opts = scene.options_struct()
        self.h = L.oracle_create(C.byref(opts), precision)
        trees = scene.tree_arrays() if hasattr(scene, "tree_arrays") else (
            [(scene.link_array(), len(scene.links), scene.self_collide)] if scene.links else [])
        self._trees = trees
        for links, count, collide in trees:
            L.oracle_add_tree(self.h, C.cast(links, C.c_void_p), count, collide)
        for b in scene.body_structs():
            L.oracle_add_body(self.h, C.cast(C.pointer(b), C.c_void_p))
        for verts, faces in getattr(scene, "meshes", []):
            v = np.ascontiguousarray(verts, dtype=np.float64)
            f = np.ascontiguousarray(faces, dtype=np.int32)
            L.oracle_add_mesh(self.h, _dp(v), len(v), f.ctypes.data_as(C.POINTER(C.c_int)), len(f))
        self._geoms = scene.geom_structs()
        for g in self._geoms:
            L.oracle_add_geom(self.h, C.cast(C.pointer(g), C.c_void_p))
        for sd in getattr(scene, "sensor_structs", lambda: [])():
            L.oracle_add_sensor(self.h, C.cast(C.pointer(sd), C.c_void_p))
        mf = tuple(getattr(scene, "magnetic_field", (0.0, 0.0, 0.0)))
        if any(mf):
            L.oracle_set_magnetic(self.h, (C.c_double * 3)(*mf))
        L.oracle_finalize(self.h, n_worlds)
        self.sensor_dim = L.oracle_sensor_dim(self.h)
        self.nq = L.oracle_nq(self.h)
        self.nv = L.oracle_nv(self.h)

    def set_state(self, qpos, qvel):
        qpos = np.ascontiguousarray(qpos, dtype=np.float64)
        qvel = np.ascontiguousarray(qvel, dtype=np.float64)
        self.L.oracle_set_state(self.h, _dp(qpos), _dp(qvel))

    def state(self):
        qpos = np.zeros((self.n, self.nq))
        qvel = np.zeros((self.n, self.nv))
        self.L.oracle_get_state(self.h, _dp(qpos), _dp(qvel))
        return qpos, qvel

    def qacc(self):
        out = np.zeros((self.n, self.nv))
        self.L.oracle_get_qacc(self.h, _dp(out))
        return out

    def set_qfrc(self, qfrc):
        """Tree::qfrc_applied of every world, [n_worlds, nv]."""
        qfrc = np.ascontiguousarray(np.broadcast_to(qfrc, (self.n, self.nv)), dtype=np.float64)
        self.L.oracle_set_qfrc(self.h, _dp(qfrc))

    def qfrc(self):
        out = np.zeros((self.n, self.nv))
        self.L.oracle_get_qfrc(self.h, _dp(out))
        return out

    def sensordata(self):
        out = np.zeros((self.n, self.sensor_dim))
        if self.sensor_dim:
            self.L.oracle_get_sensordata(self.h, _dp(out))
        return out

    def step(self, n_steps=1, threads=1):
        self.L.oracle_step(self.h, n_steps, threads)

    def detect(self, world):
        """Contacts World::detect_contacts finds at world's current state."""
        return self.L.oracle_detect(self.h, world)

    def contacts(self, world, cap=4096):
        ids = np.zeros((cap, 3), dtype=np.int32)
        geo = np.zeros((cap, 7))
        n = self.L.oracle_contacts(self.h, world, cap, ids.ctypes.data_as(C.POINTER(C.c_int)), _dp(geo))
        return ids[:n].copy(), geo[:n].copy()

    def pairs(self, world, cap=65536):
        out = np.zeros((cap, 2), dtype=np.int32)
        n = self.L.oracle_pairs(self.h, world, cap, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out[:n].copy()

    def broadphase(self, world):
        """(fat proxy bounds, swept bounds they were last updated with),
        each [n_geoms, 6] = (min xyz, max xyz)."""
        n = len(self._geoms) if hasattr(self, "_geoms") else 4096
        fat = np.zeros((n, 6))
        bound = np.zeros((n, 6))
        k = self.L.oracle_broadphase(self.h, world, _dp(fat), _dp(bound))
        return fat[:k].copy(), bound[:k].copy()

    def reinserts(self, world, reset=False):
        """Broadphase proxy (re)insertions since the last reset."""
        return int(self.L.oracle_reinserts(self.h, world, 1 if reset else 0))

    def mass_matrix(self, world):
        out = np.zeros((self.nv, self.nv))
        self.L.oracle_mass_matrix(self.h, world, _dp(out))
        return out

    def dynamics(self, world):
        """(explicit ABA qacc, CRB mass matrix, RNE bias) at the current state."""
        a = np.zeros(self.nv)
        M = np.zeros((self.nv, self.nv))
        h = np.zeros(self.nv)
        self.L.oracle_dynamics(self.h, world, _dp(a), _dp(M), _dp(h))
        return a, M, h

    def flops(self, reset=True):
        """(add/sub, mul, div, sqrt) counted on this thread since the last
        reset, for precision 2 worlds stepped with threads=1."""
        out = (C.c_ulonglong * 4)()
        self.L.oracle_flops(out, 1 if reset else 0)
        return tuple(int(x) for x in out)

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
