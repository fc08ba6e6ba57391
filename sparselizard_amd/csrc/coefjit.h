// Per-program coefficient kernels: the coefficient program of a bucket
// compiled at plan creation into straight-line device code (one thread per
// (element, Gauss point), values in registers), in place of the interpreter;
// the generated source is the prelude of build-time inlined headers
// (gen_jit_prelude.py) plus one statement per program instruction.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "kernels/device.h"

namespace slx {

// form functions per field at most for a per-element gather of its values
#define SLX_JIT_GATHER_NFF 16

// Full source of the kernel of one program (prelude included).  runs: each
// thread walks a run of Gauss points of one element (Gauss-point-varying
// buckets), and gathers the values of every field of at most
// SLX_JIT_GATHER_NFF form functions (field_nff[f] in this bucket) once for
// the run; otherwise one thread per (element, Gauss point).
std::string coef_kernel_source(const slx_instr* prog, int n_instr, bool fp64, bool runs, const std::vector<int>& field_nff);
// Compiles (or finds in the process-wide cache) the kernel of one program for
// the current device; fp64 selects FP64 coefficients.  On failure returns
// false with the compiler log in *err.
bool coef_compile(const slx_instr* prog, int n_instr, bool fp64, bool runs, const std::vector<int>& field_nff, hipFunction_t* fn,
                  std::string* err);
// Same launch contract as the interpreter's coefficient kernel.
hipError_t coef_launch(hipFunction_t fn, const DevBucket& b, const DevInputs& in, int nq, int qfast, void* ctab, hipStream_t st);

}  // namespace slx
