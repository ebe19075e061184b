"""superman_amd -- MI355X-native exact matrix permanents (Gray-code Ryser).

The compute path is libsuperman_amd.so (hand-written HIP for gfx950 behind a
C ABI, include/superman_amd.h); this package is a thin Python host layer over
it.  Importing the package does not load the library; the first call does,
and raises if the library has not been built (`python -m superman_amd.build`).
"""
from .api import (CodegenPlan, DensePlan, Permanent, apply_permutations, codegen_compile_only,  # noqa: F401
                  codegen_source, decompose, dense_c128, dense_f64, device_count, device_info, int_exact,
                  is_integral, permanent, permanent_decomposed, sparse_codegen, synchronize, ut_order)

__all__ = ["CodegenPlan", "DensePlan", "Permanent", "apply_permutations", "codegen_compile_only", "codegen_source",
           "decompose", "dense_c128", "dense_f64", "device_count", "device_info", "int_exact", "is_integral",
           "permanent", "permanent_decomposed", "sparse_codegen", "synchronize", "ut_order"]
