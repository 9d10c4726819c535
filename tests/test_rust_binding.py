"""The Rust binding shipped as source (tooling_amd/rust/newt-gpu) stays in
step with the C ABI: sys.rs declares every function include/newt_gpu.h
declares, and its #[repr(C)] records list the header's fields in order (the
ctypes mirror, whose layouts test_abi.py checks against the C compiler).
No Rust toolchain exists here, so the crate itself is not compiled.
"""
import os
import re

from tooling_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SYS = os.path.join(ROOT, "tooling_amd", "rust", "newt-gpu", "src", "sys.rs")


HEADER = os.path.join(ROOT, "include", "newt_gpu.h")


def _header_symbols():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(newt_[a-z_0-9]+)\s*\(", text)))


def test_sys_rs_declares_every_header_function():
    text = open(SYS).read()
    rust = sorted(set(re.findall(r"pub fn (newt_[a-z_0-9]+)\(", text)))
    assert rust == _header_symbols()


def test_sys_rs_records_match_the_header_fields():
    text = open(SYS).read()
    records = {"NewtOptions": abi.Options, "NewtLinkDesc": abi.LinkDesc, "NewtGeomDesc": abi.GeomDesc,
               "NewtBodyDesc": abi.BodyDesc, "NewtDeviceState": abi.DeviceState,
               "NewtBatchOptions": abi.BatchOptions, "NewtSensorDesc": abi.SensorDesc}
    for name, py in records.items():
        m = re.search(r"pub struct " + name + r" \{(.*?)\n\}", text, re.S)
        assert m, name
        fields = re.findall(r"pub ([a-z_0-9]+):", m.group(1))
        assert fields == [f for f, _ in py._fields_], name
