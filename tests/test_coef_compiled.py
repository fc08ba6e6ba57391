"""Compiled coefficient programs (slx_plan_coef_mode) against the
interpreter and the oracle.

Each bucket's coefficient program can run as its own straight-line kernel
compiled at plan creation; it performs the interpreter's arithmetic in the same
order and precision.  The coefficient tables and the assembled CSR of both
modes are compared, and the compiled mode is checked against the CPU oracle
like every other path."""
import numpy as np
import pytest

import sparselizard_amd as sa
from oracle import py_oracle

pytestmark = pytest.mark.gpu

CASES = [("elasticity3d", "n=3"), ("newton2d", "n=6"), ("projection", "elem=hex;n=3;order=2"),
         ("projection", "elem=quad;n=4;order=3"), ("hppoisson3d", "n=2"), ("quadbatch", "instances=3;nmin=2;nmax=5"),
         ("prism3d", "n=3;order=2")]


def relerr(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("kind,spec", CASES)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_compiled_matches_interpreter(ctx, kind, spec, precision):
    print("case compiled coefficients", kind, spec, precision, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    info = prob.info()
    plan = prob.plan()
    assert plan.coef_mode("interpreter") == 0
    prob.generate(precision)
    c_int = plan.coefficients()
    m_int = plan.csr(info["ndofs"])
    assert plan.coef_mode("compiled") >= 2
    prob.generate(precision)
    c_jit = plan.coefficients()
    m_jit = plan.csr(info["ndofs"])
    same = np.array_equal(c_int, c_jit)
    print("bitwise equal coefficients:", same, "rel", relerr(c_jit, c_int), flush=True)
    tol = 1e-15 if precision == "fp64" else 1e-7
    assert relerr(c_jit, c_int) <= tol
    for name in ("Aval", "Dval", "rhs"):
        a, b = getattr(m_jit, name), getattr(m_int, name)
        if b.size:
            assert relerr(a, b) <= tol, name
    plan.coef_mode("auto")


@pytest.mark.parametrize("kind,spec,precision,tol", [("elasticity3d", "n=3", "f16split", 1e-6), ("newton2d", "n=6", "fp64", 1e-12),
                                                     ("hppoisson3d", "n=2", "f16split", 1e-6)])
def test_compiled_matches_oracle(ctx, kind, spec, precision, tol):
    print("case compiled vs oracle", kind, spec, precision, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    info = prob.info()
    plan = prob.plan()
    plan.coef_mode("compiled")
    got = prob.generate(precision).csr(info["ndofs"])
    ref = py_oracle.assemble(prob.desc_ptr(), prob.numeric(), info["ndofs"], info["nfree"])
    for name in ("Arow", "Acol", "Drow", "Dcol"):
        assert np.array_equal(getattr(got, name), getattr(ref, name)), name
    for name in ("Aval", "Dval", "rhs"):
        b = getattr(ref, name)
        if b.size:
            assert relerr(getattr(got, name), b) <= tol, name


def test_coef_mode_argument_error(ctx):
    prob = sa.Problem(ctx, "poisson2d", "n=4")
    with pytest.raises(KeyError):
        prob.plan().coef_mode("jit")
