"""Model / Data / step over the C ABI — the host-side mirror of newt's World.

newt keeps model and state in one `World` and advances it with
`World::step(&mut self)` (newt/src/world.rs:1459).  Here a :class:`Model`
is the immutable half (options, tree, geoms; compiled into device tables)
and a :class:`Data` holds the state of N worlds on the GPU; ``step(model,
data, n)`` advances all of them by ``n`` World::step calls.  Errors raise
:class:`tooling_amd.abi.NewtError` where newt panics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .scenes import Scene


class Model:
    """`contact_capacity`: most contacts per world (newt_model_set_contact_capacity);
    a world that finds more freezes with abi.ERR_CAPACITY."""

    def __init__(self, scene: Scene, contact_capacity: int | None = None):
        self.lib = abi.load_library()
        self.scene = scene
        h = C.c_void_p()
        opts = scene.options_struct()
        abi.check(self.lib.newt_model_create(C.byref(opts), C.byref(h)))
        self.handle = h
        for links, count, collide in scene.tree_arrays():
            abi.check(self.lib.newt_model_add_tree(h, links, count, collide, None))
        for b in scene.body_structs():
            abi.check(self.lib.newt_model_add_body(h, C.byref(b), None))
        for verts, faces in getattr(scene, "meshes", []):
            v = np.ascontiguousarray(verts, dtype=np.float64)
            f = np.ascontiguousarray(faces, dtype=np.int32)
            abi.check(self.lib.newt_model_add_mesh(h, v.ctypes.data_as(C.POINTER(C.c_double)), len(v),
                                                   f.ctypes.data_as(C.POINTER(C.c_int32)), len(f), None))
        for g in scene.geom_structs():
            abi.check(self.lib.newt_model_add_geom(h, C.byref(g), None))
        if contact_capacity is not None:
            abi.check(self.lib.newt_model_set_contact_capacity(h, int(contact_capacity)))
        for sd in scene.sensor_structs():
            abi.check(self.lib.newt_model_add_sensor(h, C.byref(sd), None))
        if any(scene.magnetic_field):
            mf = (C.c_double * 3)(*scene.magnetic_field)
            abi.check(self.lib.newt_model_set_magnetic_field(h, mf))
        abi.check(self.lib.newt_model_finalize(h))
        dim = C.c_int32()
        abi.check(self.lib.newt_model_sensor_dim(h, C.byref(dim)))
        self.sensor_dim = dim.value
        nq, nv, nl, ng = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        abi.check(self.lib.newt_model_dims(h, C.byref(nq), C.byref(nv), C.byref(nl), C.byref(ng)))
        self.nq, self.nv, self.n_links, self.n_geoms = nq.value, nv.value, nl.value, ng.value

    def close(self):
        if self.handle:
            self.lib.newt_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Data:
    def __init__(self, model: Model, n_worlds: int, precision: int = abi.PRECISION_FP64):
        self.model = model
        self.lib = model.lib
        self.n_worlds = n_worlds
        self.precision = precision
        h = C.c_void_p()
        abi.check(self.lib.newt_data_create(model.handle, n_worlds, precision, C.byref(h)))
        self.handle = h

    @staticmethod
    def _dp(a):
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def set_state(self, qpos, qvel):
        qpos = np.ascontiguousarray(qpos, dtype=np.float64).reshape(self.n_worlds, self.model.nq)
        qvel = np.ascontiguousarray(qvel, dtype=np.float64).reshape(self.n_worlds, self.model.nv)
        abi.check(self.lib.newt_data_set_state(self.handle, self._dp(qpos), self._dp(qvel)))

    def set_state_async(self, qpos, qvel, stream=None):
        """set_state ordered on `stream` (a hipStream_t handle or None); the
        arrays are staged before the call returns."""
        qpos = np.ascontiguousarray(qpos, dtype=np.float64).reshape(self.n_worlds, self.model.nq)
        qvel = np.ascontiguousarray(qvel, dtype=np.float64).reshape(self.n_worlds, self.model.nv)
        abi.check(self.lib.newt_data_set_state_async(self.handle, self._dp(qpos), self._dp(qvel),
                                                     C.c_void_p(stream) if stream else None))

    def set_worlds_async(self, first, qpos, qvel, stream=None):
        """Start len(qpos) new worlds in slots [first, first+len(qpos)),
        ordered on `stream`; the arrays are staged before the call returns."""
        qpos = np.ascontiguousarray(qpos, dtype=np.float64).reshape(-1, self.model.nq)
        qvel = np.ascontiguousarray(qvel, dtype=np.float64).reshape(-1, self.model.nv)
        if qpos.shape[0] != qvel.shape[0]:
            raise ValueError("qpos and qvel hold different world counts")
        abi.check(self.lib.newt_data_set_worlds_async(self.handle, int(first), qpos.shape[0], self._dp(qpos),
                                                      self._dp(qvel), C.c_void_p(stream) if stream else None))

    def park_worlds(self, first, count, stream=None):
        """Stop stepping slots [first, first+count) until they are reused."""
        abi.check(self.lib.newt_data_park_worlds(self.handle, int(first), int(count),
                                                 C.c_void_p(stream) if stream else None))

    def worlds(self, first, count):
        qpos = np.zeros((count, self.model.nq))
        qvel = np.zeros((count, self.model.nv))
        abi.check(self.lib.newt_data_get_worlds(self.handle, int(first), int(count), self._dp(qpos),
                                                self._dp(qvel)))
        return qpos, qvel

    def device_bytes(self):
        out = C.c_uint64()
        abi.check(self.lib.newt_data_device_bytes(self.handle, C.byref(out)))
        return out.value

    def state(self):
        qpos = np.zeros((self.n_worlds, self.model.nq))
        qvel = np.zeros((self.n_worlds, self.model.nv))
        abi.check(self.lib.newt_data_get_state(self.handle, self._dp(qpos), self._dp(qvel)))
        return qpos, qvel

    def set_qfrc(self, qfrc):
        """Tree::qfrc_applied of every world, [n_worlds, nv] (broadcast)."""
        qfrc = np.ascontiguousarray(np.broadcast_to(qfrc, (self.n_worlds, self.model.nv)), dtype=np.float64)
        abi.check(self.lib.newt_data_set_qfrc_applied(self.handle, self._dp(qfrc)))

    def qfrc(self):
        out = np.zeros((self.n_worlds, self.model.nv))
        abi.check(self.lib.newt_data_get_qfrc_applied(self.handle, self._dp(out)))
        return out

    def qacc(self):
        out = np.zeros((self.n_worlds, self.model.nv))
        abi.check(self.lib.newt_data_get_qacc(self.handle, self._dp(out)))
        return out

    def sensordata(self):
        """World::sensordata after the last step, [n_worlds, sensor_dim]."""
        out = np.zeros((self.n_worlds, max(self.model.sensor_dim, 0)))
        if self.model.sensor_dim:
            abi.check(self.lib.newt_data_get_sensordata(self.handle, self._dp(out)))
        return out

    def faults(self):
        out = np.zeros(self.n_worlds, dtype=np.int32)
        abi.check(self.lib.newt_data_get_faults(self.handle, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def contacts(self):
        """Per world: list of (geom_a, geom_b, slot) and an array of
        (pos xyz, normal xyz, penetration) rows, in narrow-phase order."""
        cap = C.c_int32()
        abi.check(self.lib.newt_data_contact_capacity(self.handle, C.byref(cap)))
        cap = cap.value
        counts = np.zeros(self.n_worlds, dtype=np.int32)
        ids = np.zeros((self.n_worlds, cap, 3), dtype=np.int32)
        geo = np.zeros((self.n_worlds, cap, 7))
        abi.check(self.lib.newt_data_get_contacts(
            self.handle, counts.ctypes.data_as(C.POINTER(C.c_int32)),
            ids.ctypes.data_as(C.POINTER(C.c_int32)), self._dp(geo)))
        return counts, ids, geo

    def pairs(self):
        cap = C.c_int32()
        abi.check(self.lib.newt_data_pair_capacity(self.handle, C.byref(cap)))
        cap = cap.value
        counts = np.zeros(self.n_worlds, dtype=np.int32)
        pairs = np.zeros((self.n_worlds, cap, 2), dtype=np.int32)
        abi.check(self.lib.newt_data_get_pairs(
            self.handle, counts.ctypes.data_as(C.POINTER(C.c_int32)),
            pairs.ctypes.data_as(C.POINTER(C.c_int32))))
        return counts, pairs

    def synchronize(self):
        abi.check(self.lib.newt_data_synchronize(self.handle))

    def _mask(self, mask):
        if mask is None:
            return None, None
        m = np.ascontiguousarray(np.asarray(mask, dtype=bool).astype(np.uint8).reshape(self.n_worlds))
        return m, m.ctypes.data_as(C.POINTER(C.c_uint8))

    def reset(self, mask=None, stream=None):
        """Masked worlds (all when None) back to the model's default state."""
        keep, mp = self._mask(mask)
        abi.check(self.lib.newt_data_reset(self.handle, mp, C.c_void_p(stream) if stream else None))
        self.synchronize()

    def reset_to_keyframe(self, qpos, qvel, mask=None, stream=None):
        """Masked worlds take the keyframe (qpos [nq], qvel [nv])."""
        keep, mp = self._mask(mask)
        q = np.ascontiguousarray(qpos, dtype=np.float64).reshape(self.model.nq)
        v = np.ascontiguousarray(qvel, dtype=np.float64).reshape(self.model.nv)
        abi.check(self.lib.newt_data_reset_to_keyframe(self.handle, mp, self._dp(q), self._dp(v),
                                                       C.c_void_p(stream) if stream else None))
        self.synchronize()

    def snapshot(self):
        return Snapshot(self)

    def restore(self, snap, stream=None):
        abi.check(self.lib.newt_data_restore(self.handle, snap.handle, C.c_void_p(stream) if stream else None))
        self.synchronize()

    def close(self):
        if self.handle:
            self.lib.newt_data_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Snapshot:
    """A device copy of a Data's state (newt_data_snapshot)."""

    def __init__(self, data: Data):
        self.lib = data.lib
        h = C.c_void_p()
        abi.check(self.lib.newt_data_snapshot(data.handle, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.newt_snapshot_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step(model: Model, data: Data, n_steps: int = 1, stream=None):
    """Enqueue n_steps World::step calls for every world of `data`."""
    s = C.c_void_p(stream) if stream is not None else None
    abi.check(model.lib.newt_step(model.handle, data.handle, int(n_steps), s))


class Batch:
    """Several Data of different models stepped together (newt_batch): one
    launch per kernel instantiation, replayed from a HIP graph."""

    def __init__(self, datas, order=1, wpb_cap=(0, 0, 0), use_graph=1, branches=1):
        self.datas = list(datas)
        self.lib = self.datas[0].lib
        arr = (C.c_void_p * len(self.datas))(*[d.handle.value for d in self.datas])
        opt = abi.BatchOptions(order=order, wpb_cap=(C.c_int32 * 3)(*wpb_cap), use_graph=use_graph,
                               branches=branches)
        h = C.c_void_p()
        abi.check(self.lib.newt_batch_create(arr, len(self.datas), C.byref(opt), C.byref(h)))
        self.handle = h

    def step(self, n_steps=1, stream=None):
        s = C.c_void_p(stream) if stream is not None else None
        abi.check(self.lib.newt_batch_step(self.handle, int(n_steps), s))

    def launches(self):
        n = C.c_int32()
        abi.check(self.lib.newt_batch_launches(self.handle, C.byref(n)))
        return n.value

    def close(self):
        if self.handle:
            self.lib.newt_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
