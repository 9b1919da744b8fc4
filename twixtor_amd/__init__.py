"""twixtor_amd: MI355X (gfx950) HIP backend of twixtor's frame-synthesis hot path.

The product is libtwx.so (C ABI in include/twx.h). This package is the host-side mirror of the
reference's device backend interface (plugins/retime-engine/src/device/exec/backend.rs:27-46) over
that ABI, plus the `vfx apply retime` parameter mapping and the synthetic clip generators of the
benchmark workloads.
"""
from .params import Retime, RetimeParams, Selection  # noqa: F401
