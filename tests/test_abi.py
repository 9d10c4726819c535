"""The C ABI library builds, loads and exports every declared symbol."""
import ctypes
import os
import re

from tooling_amd import abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "newt_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(newt_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = declared_symbols()
    assert "newt_step" in names and "newt_model_create" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert sorted(abi.SIGNATURES) == declared_symbols()


def test_struct_layouts_match_the_header():
    """Size and every field offset of the ctypes records against the C
    compiler's layout of include/newt_gpu.h."""
    import shutil
    import subprocess
    import tempfile
    import pytest
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"newt_options": abi.Options, "newt_link_desc": abi.LinkDesc, "newt_geom_desc": abi.GeomDesc,
               "newt_body_desc": abi.BodyDesc, "newt_device_state": abi.DeviceState,
               "newt_batch_options": abi.BatchOptions, "newt_sensor_desc": abi.SensorDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof() %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run([cc, "-std=c99", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line.strip():
            cname, field, value = line.split()
            got[(cname, field)] = int(value)
    for cname, py in structs.items():
        assert got[(cname, "sizeof()")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
