"""gpurify_amd: MI355X (gfx950) quasi-static capacitance engine.

The compute path is libgpurify.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/gpurify.h).  This package holds its ctypes binding and
the seeded synthetic layouts used by the tests and the benchmark.  Importing
a solve without the built library raises; there is no CPU fallback.
"""
from ._lib import (  # noqa: F401
    Context, GpError, default_opts, device_ready, load, m2l_factors, m2l_operator, mesh_groups, parse_qui,
    solve_resistance, version,
    GP_PREC_FP64, GP_PREC_MIXED, GP_METHOD_AUTO, GP_METHOD_DENSE, GP_METHOD_FMM,
    GP_M2L_FP64, GP_M2L_FP32, GP_M2L_BF16, GP_M2L_FP64_VALU, GP_M2L_BF16_VALU, GP_BASIS_FP32, GP_BASIS_FP64, GP_NEAR_FP32, GP_NEAR_FP16,
    GP_GROUND_NONE, GP_GROUND_IMAGE,
    LIB_PATH,
)
from . import workloads  # noqa: F401
