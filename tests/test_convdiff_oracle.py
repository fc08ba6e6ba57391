"""The convection-diffusion workload (a non-symmetric system for the GMRES
path), pinned on the oracle by closed forms.

With b constant (div b = 0), the convection matrix C_ij = int (b . grad phi_j)
phi_i satisfies C + C^T = boundary terms, which vanish between functions that
are zero on the boundary (every free DOF here).  So on the free DOFs the
symmetric part of A is exactly the Poisson stiffness matrix of the same mesh,
and A 1 + D 1 = 0 at order 1 (constants are in the kernel of both
operators)."""
import numpy as np
import scipy.sparse as sp

import sparselizard_amd as sa
from oracle import py_oracle


def _assemble(kind, spec):
    p = sa.Problem(None, kind, spec)
    info = p.info()
    o = py_oracle.assemble(p.desc_ptr(), p.numeric(), info["ndofs"], info["nfree"])
    nf, nc = info["nfree"], info["ndofs"] - info["nfree"]
    A = sp.csr_matrix((o.Aval, o.Acol, o.Arow), shape=(nf, nf))
    D = sp.csr_matrix((o.Dval, o.Dcol, o.Drow), shape=(nf, nc))
    return A, D, o.rhs


def test_convdiff_symmetric_part_is_poisson_stiffness():
    for order in (1, 2):
        A, D, b = _assemble("convdiff2d", "n=12;order=%d;bx=20;by=10" % order)
        K, DK, bK = _assemble("poisson2d", "n=12;order=%d" % order)
        S = 0.5 * (A + A.T)
        assert abs(S - K).max() <= 1e-12 * abs(K).max()
        # the convection part is non-trivial and skew on the free DOFs
        C = A - K
        assert abs(C).max() > 1e-2 * abs(K).max()
        assert abs(C + C.T).max() <= 1e-12 * abs(K).max()
        assert np.allclose(b, bK, rtol=0, atol=1e-14)


def test_convdiff_constants_in_kernel():
    # order 1: the all-ones DOF vector is the constant function (the
    # hierarchical higher-order functions are not nodal)
    A, D, b = _assemble("convdiff2d", "n=10;order=1;bx=-7;by=3")
    r = A @ np.ones(A.shape[0]) + D @ np.ones(D.shape[1])
    assert abs(r).max() <= 1e-11 * abs(A).max()
