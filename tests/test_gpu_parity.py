"""GPU parity: the HIP assembly path against the CPU FP64 oracle.

Pattern (rowptr / column indices of A and D) must be bit-exact; values match
to the stated relative Frobenius tolerance.  Every case runs through the C ABI
(slx_problem_generate -> slx_plan_create / slx_assemble / slx_csr_copyout).
"""
import numpy as np
import pytest

import sparselizard_amd as sa
from oracle import py_oracle

pytestmark = pytest.mark.gpu

# FP64 kernels reorder sums relative to the oracle (contraction order, FMA):
# only rounding differences remain.
TOL_FP64 = 1e-12
# FP32 element kernel, FP64 reduction.
TOL_FP32 = 2e-5
# FP16 hi/lo split MFMA with FP32 accumulation (north-star gate 1e-6).
TOL_SPLIT = 1e-6


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def run_case(ctx, kind, spec, precision, tol):
    print("case", kind, spec, precision, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    info = prob.info()
    plan = prob.generate(precision)
    got = plan.csr(info["ndofs"])
    ref = py_oracle.assemble(prob.desc_ptr(), prob.numeric(), info["ndofs"], info["nfree"])
    for name in ("Arow", "Acol", "Drow", "Dcol"):
        a, b = getattr(got, name), getattr(ref, name)
        assert a.shape == b.shape and np.array_equal(a, b), "pattern mismatch in " + name
    errs = {}
    for name in ("Aval", "Dval", "rhs"):
        a, b = getattr(got, name), getattr(ref, name)
        if b.size:
            errs[name] = relerr(a, b)
            assert errs[name] <= tol, "%s rel err %.3e > %.1e" % (name, errs[name], tol)
    return prob, plan, got, errs


@pytest.mark.parametrize("n", [2, 9, 40])
def test_poisson2d_fp64(ctx, n):
    run_case(ctx, "poisson2d", "n=%d" % n, "fp64", TOL_FP64)


def test_poisson2d_p3_orientation_buckets(ctx):
    # order 3 on triangles: orientation-dependent face functions, one bucket per total orientation
    run_case(ctx, "poisson2d", "n=6;order=3", "fp64", TOL_FP64)


@pytest.mark.parametrize("precision,tol", [("fp64", TOL_FP64), ("fp32", TOL_FP32)])
def test_elasticity3d_small(ctx, precision, tol):
    run_case(ctx, "elasticity3d", "n=3", precision, tol)


def test_quadbatch_small_fp64(ctx):
    run_case(ctx, "quadbatch", "instances=3;nmin=2;nmax=5", "fp64", TOL_FP64)


def test_mfma_operand_layout(ctx):
    # A[m][k] = m + k/64, B selects rows n and n+16 of column n: D[m][n] = A[m][n] + A[m][n+16]
    from sparselizard_amd import _lib as L
    d = L.mfma_layout_probe(ctx)
    m = np.arange(16)[:, None]
    n = np.arange(16)[None, :]
    expect = (m + n / 64.0) + (m + (n + 16) / 64.0)
    assert np.array_equal(d, expect.astype(np.float32))


@pytest.mark.parametrize("kind,spec", [("poisson2d", "n=9"), ("poisson2d", "n=7;order=2"), ("elasticity3d", "n=3"),
                                       ("elasticity3d", "n=4;order=1"), ("quadbatch", "instances=3;nmin=2;nmax=5")])
def test_f16split_mfma_matches_oracle(ctx, kind, spec):
    prob, plan, got, errs = run_case(ctx, kind, spec, "f16split", TOL_SPLIT)
    assert plan.kernels()["mfma_buckets"] >= 1, "MFMA kernel not used"
    print("errors", errs)


@pytest.mark.parametrize("precision", ["fp64", "f16split"])
def test_deterministic_repeat(ctx, precision):
    prob = sa.Problem(ctx, "elasticity3d", "n=3")
    info = prob.info()
    a = prob.generate(precision).csr(info["ndofs"])
    b = prob.generate(precision).csr(info["ndofs"])
    assert np.array_equal(a.Aval, b.Aval) and np.array_equal(a.Dval, b.Dval) and np.array_equal(a.rhs, b.rhs)
