"""slx_plan_solve_noalloc (workspace in the library's fixed per-GPU array)
against slx_plan_solve on the same assembled systems: identical solutions,
counts and residuals (both are deterministic and run the same arithmetic),
on two calls in a row, the second on a smaller system (the array then holds
the first call's contents), for the CG and the GMRES inner solves; a system
whose workspace exceeds the array is rejected before any device work."""
import numpy as np
import pytest

import sparselizard_amd as sa
from sparselizard_amd import _lib as L

pytestmark = pytest.mark.gpu


def assembled(ctx, kind, spec, precision="fp64"):
    print("case noalloc", kind, spec, precision, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    info = prob.info()
    plan = prob.generate(precision)
    return prob, plan, info["ndofs"]


@pytest.mark.parametrize("kind,big,small,method", [
    ("poisson2d", "n=80", "n=30", "cg"),
    ("elasticity3d", "n=4", "n=2", "auto"),       # symmetry check in the array's header, then CG
    ("convdiff2d", "n=40", "n=16", "auto"),       # non-symmetric: GMRES basis in the array
])
def test_noalloc_solve_equals_solve(ctx, kind, big, small, method):
    pb, plan_b, nb = assembled(ctx, kind, big)
    ps, plan_s, ns = assembled(ctx, kind, small)
    kw = dict(rel_tol=1e-10, max_outer=20, max_inner=200, method=method)
    xb, ib = plan_b.solve(nb, **kw)
    xs, is_ = plan_s.solve(ns, **kw)
    # two no-allocation calls in a row, the second on the smaller system
    xb2, ib2 = plan_b.solve_noalloc(nb, **kw)
    xs2, is2 = plan_s.solve_noalloc(ns, **kw)
    print("big", ib, "small", is_, flush=True)
    assert np.array_equal(xb, xb2) and ib == ib2
    assert np.array_equal(xs, xs2) and is_ == is2
    assert ib["rel_residual"] <= 1e-10 and is_["rel_residual"] <= 1e-10


def test_noalloc_solve_rejects_oversized_workspace(ctx):
    # GMRES with 100 Krylov vectors of 3.6e5 floats: 146 MB > 128 MiB
    prob, plan, nd = assembled(ctx, "poisson2d", "n=600")
    assert plan.shape()["nfree"] > 350000
    with pytest.raises(L.SlxError, match="exceeds"):
        plan.solve_noalloc(nd, rel_tol=1e-8, max_outer=2, max_inner=100, method="gmres")
    # the plan still solves through the allocating entry point afterwards
    x, info = plan.solve(nd, rel_tol=1e-8, max_outer=3, max_inner=100, method="cg")
    assert info["outer"] >= 1
