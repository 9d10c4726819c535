"""MI355X-native batched step of newt's rigid-body World.

The HIP library (tooling_amd/_build/libnewt_gpu.so) exposes the C ABI of
include/newt_gpu.h; :mod:`tooling_amd.newt` mirrors it as Model / Data /
step, and :mod:`tooling_amd.scenes` builds the benchmark scenes.
"""
from .abi import LinkDesc, GeomDesc, Options, load_library  # noqa: F401
