"""MI355X-native weak-form assembly (sparselizard's formulation::generate path).

The product is the C++/HIP library sparselizard_amd/lib/libslx.so behind the C
ABI of include/slx.h; this package only binds it for tests and benchmarks.
"""
from ._lib import Context, Problem, Plan, CSR, SlxError, PRECISIONS, lib, LIB_PATH  # noqa: F401
