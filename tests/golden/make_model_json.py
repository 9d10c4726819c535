"""Writes tests/golden/models/quadruped18.json: the 18-DoF quadruped of
tooling_amd.scenes.quadruped() in newt's JSON scene format
(docs/model-format.md), with explicit diagonal inertias and a ground plane,
so the loader can be checked against the programmatic scene.

    python tests/golden/make_model_json.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tooling_amd import abi, scenes  # noqa: E402

KIND = {abi.JOINT_FREE: "free", abi.JOINT_FIXED: "fixed", abi.JOINT_HINGE: "hinge", abi.JOINT_SLIDE: "slide"}
SHAPE = {abi.SHAPE_PLANE: "plane", abi.SHAPE_SPHERE: "sphere", abi.SHAPE_BOX: "box", abi.SHAPE_CAPSULE: "capsule"}


def main():
    sc = scenes.quadruped()
    links = []
    for i, l in enumerate(sc.links):
        j = {"kind": KIND[l["joint"]]}
        if l["joint"] in (abi.JOINT_HINGE, abi.JOINT_SLIDE):
            j["axis"] = list(l["axis"])
        d = {"name": f"l{i}", "joint": j,
             "joint_offset_in_parent": {"position": list(l["parent_pos"]), "orientation": list(l["parent_quat"])},
             "joint_offset_in_child": {"position": list(l["child_pos"])},
             "mass": l["mass"], "inertia": {"kind": "diag", "values": [l["inertia"][0], l["inertia"][4],
                                                                          l["inertia"][8]]}}
        if i > 0:
            d["parent"] = f"l{l['parent']}"
            del d["joint_offset_in_parent"]["orientation"]
        links.append(d)
    geoms = []
    for k, g in enumerate(sc.geoms):
        shape = {"kind": SHAPE[g["shape"]]}
        if g["shape"] == abi.SHAPE_SPHERE:
            shape["radius"] = g["size"][0]
        elif g["shape"] == abi.SHAPE_BOX:
            shape["half_extents"] = list(g["size"])
        elif g["shape"] == abi.SHAPE_CAPSULE:
            shape["radius"], shape["half_height"] = g["size"][0], g["size"][1]
        attach = {"kind": "static"} if g["link"] < 0 else {"kind": "link", "tree": "quadruped", "link": f"l{g['link']}"}
        geoms.append({"name": f"g{k}", "shape": shape, "attach": attach, "local_offset": list(g["pos"]),
                      "local_orientation": list(g["quat"]), "friction": g["friction"]})
    o = sc.options
    doc = {"version": "1", "gravity": list(o["gravity"]), "timestep": o["dt"], "integrator": "Euler",
           "solver": {"mode": "pgs", "iterations": o["iterations"], "cone": "pyramidal"},
           "trees": [{"name": "quadruped", "self_collide": False, "links": links}], "geoms": geoms}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "quadruped18.json")
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
