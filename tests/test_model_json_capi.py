"""newt_model_from_json (the C++ loader behind the C ABI) against the Python
loader (tooling_amd/model_json.py) it mirrors: the same link, geom and body
descriptions field for field, bit for bit, on newt's model files and the
quadruped; the same error paths on strictness violations.  The model is
loaded without finalize, so no device is needed.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from tooling_amd import abi, model_json

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


def _load_c(text):
    lib = abi.load_library()
    h = C.c_void_p()
    st = lib.newt_model_from_json(text.encode(), 0, C.byref(h))
    if st != abi.OK:
        return None, lib.newt_last_error().decode()
    return h, None


def _fields(struct):
    out = {}
    for name, _ in struct._fields_:
        v = getattr(struct, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def _c_descs(h):
    lib = abi.load_library()
    nt, nb, ng = C.c_int32(), C.c_int32(), C.c_int32()
    abi.check(lib.newt_model_counts(h, C.byref(nt), C.byref(nb), C.byref(ng)))
    opts = abi.Options()
    abi.check(lib.newt_model_get_options(h, C.byref(opts)))
    links, sc = [], 0
    if nt.value:
        n, s = C.c_int32(), C.c_int32()
        abi.check(lib.newt_model_get_tree(h, 0, None, 0, C.byref(n), C.byref(s)))
        arr = (abi.LinkDesc * n.value)()
        abi.check(lib.newt_model_get_tree(h, 0, arr, n.value, C.byref(n), C.byref(s)))
        links, sc = [_fields(arr[i]) for i in range(n.value)], s.value
    geoms = []
    for i in range(ng.value):
        g = abi.GeomDesc()
        abi.check(lib.newt_model_get_geom(h, i, C.byref(g)))
        geoms.append(_fields(g))
    bodies = []
    for i in range(nb.value):
        b = abi.BodyDesc()
        abi.check(lib.newt_model_get_body(h, i, C.byref(b)))
        bodies.append(_fields(b))
    lib.newt_model_destroy(h)
    return _fields(opts), links, sc, geoms, bodies


def _py_descs(text):
    sc = model_json.load(text)
    links = [_fields(sc.link_array()[i]) for i in range(len(sc.links))] if sc.links else []
    return (_fields(sc.options_struct()), links, sc.self_collide, [_fields(g) for g in sc.geom_structs()],
            [_fields(b) for b in sc.body_structs()])


@pytest.mark.parametrize("name", ["pendulum", "stack", "quadruped18"])
def test_c_loader_matches_python_loader(name):
    text = open(os.path.join(MODELS, f"{name}.json")).read()
    h, err = _load_c(text)
    assert h is not None, err
    got, want = _c_descs(h), _py_descs(text)
    labels = ("options", "links", "self_collide", "geoms", "bodies")
    for label, g, w in zip(labels, got, want):
        assert g == w, f"{name}: {label} differ"


def _doc(name="pendulum"):
    with open(os.path.join(MODELS, f"{name}.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("mutate", [
    lambda d: d["trees"][0]["links"][1]["joint"].update(dampign=0.1),
    lambda d: d["trees"][0]["links"][1]["joint"].update(axis=[0, 0, 0]),
    lambda d: d["trees"][0]["links"][1]["joint"].update(kind="ball"),
    lambda d: d["trees"][0]["links"][2].update(parent="nowhere"),
    lambda d: d["trees"][0]["links"][1]["joint"].update(range=[1.0, -1.0]),
    lambda d: d["trees"][0]["links"][1].update(mass=-1.0),
    lambda d: d.update(integrator="Verlet"),
    lambda d: d.update(timestep=0),
    lambda d: d.update(solver={"mode": "pgs", "iterations": 0}),
    lambda d: d.update(geoms=[{"shape": {"kind": "sphere", "radius": 0.1}}]),
    lambda d: d.update(geoms=[{"shape": {"kind": "cylinder", "radius": 0.1, "half_height": 0.2},
                               "attach": {"kind": "link", "tree": "pendulum", "link": "upper"}}]),
    lambda d: d.update(trees=d["trees"] * 2),
    lambda d: d.update(tendons=[{"name": "t"}]),
])
def test_c_loader_errors_name_the_same_path(mutate):
    d = _doc()
    mutate(d)
    text = json.dumps(d)
    with pytest.raises(model_json.ModelError) as e:
        model_json.load(text)
    h, err = _load_c(text)
    assert h is None
    assert err.split(": ")[0] == e.value.path, (err, e.value.path)


def test_c_loader_rejects_bad_json():
    for text in ('{"timestep": 0.01,}', '{"a": 1, "a": 2}', '[1, 2', '{"gravity": [0, 0, -9.81]} x'):
        h, err = _load_c(text)
        assert h is None and err.startswith("<json>"), (text, err)
