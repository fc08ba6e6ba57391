"""Curved (order-2) geometry and the elasticity-membrane-3d golden.

  examples/elasticity-membrane-3d/main.cpp:50
      -1.24113e-10 < compz(u).integrate(vol, u, 5) < -1.24111e-10
  for the h1xyz displacement of order 3 on the example's disk of 32 curved
  27-node hexahedra (disk.msh, copied as data into tests/golden/), clamped on
  surface 2, E = 150e9, nu = 0.3, volume force (0, 0, -10).

The integral of u_z at integration order 5 is rhs_5 . x, where rhs_5 is the
right-hand side of the same formulation with integral(vol, (0,0,-1)*tf(u), -3)
(problem option measure=1; same fields, hence the same DOF numbering).  The
mesh moves by u in the reference's line (integrate(vol, u, 5)); at a
displacement of 1e-10 m that changes the integral far below the window.

Geometry functions of order 2 are pinned by two independent constructions:
the product's closed forms (basis.cpp lagrange2_tables) against the oracle's
inverse Vandermonde matrices of the same node sets (oracle.cpp lagrange2),
plus partition of unity.  The solve x_A = A^-1 (b_A - D x_D) is scipy's
(CPU) or slx_plan_solve (GPU).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import sparselizard_amd as sa
from sparselizard_amd import _lib as L
from oracle import py_oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DISK = "file=" + os.path.join(GOLDEN, "membrane3d_disk.msh")
LO, HI = -1.24113e-10, -1.24111e-10
# element type: (node count, reference-element sampler)
TYPES = {1: 3, 2: 6, 3: 9, 4: 10, 5: 27, 6: 18}


def sample(t, n, rng):
    u = rng.uniform(0, 1, (n, 3))
    if t in (1, 3, 5):                      # [-1, 1]^d
        p = 2 * u - 1
        if t == 1:
            p[:, 1:] = 0
        if t == 3:
            p[:, 2] = 0
        return p
    if t == 2:                              # triangle
        m = u[:, 0] + u[:, 1] > 1
        u[m, :2] = 1 - u[m, :2]
        u[:, 2] = 0
        return u
    if t == 4:                              # tetrahedron (rejection)
        return u[u.sum(axis=1) <= 1]
    m = u[:, 0] + u[:, 1] > 1               # prism: triangle x [-1, 1]
    u[m, :2] = 1 - u[m, :2]
    u[:, 2] = 2 * u[:, 2] - 1
    return u


@pytest.mark.parametrize("t", sorted(TYPES))
def test_order2_geometry_functions(t):
    pts = sample(t, 40, np.random.default_rng(t))
    got = L.geometry_tables(t, TYPES[t], pts)
    ref = py_oracle.lagrange2(t, pts)
    assert ref is not None and ref.shape == got.shape
    assert np.abs(got - ref).max() <= 1e-12
    assert np.allclose(got[0].sum(axis=0), 1.0, atol=1e-13)   # partition of unity
    assert np.allclose(got[1:].sum(axis=1), 0.0, atol=1e-12)  # its derivatives


def test_order2_geometry_detects_a_wrong_node_order():
    # a plausible mistake (two edge midpoints of the hexahedron swapped) breaks
    # the agreement: the oracle's functions follow its own node table
    pts = sample(5, 20, np.random.default_rng(0))
    got = L.geometry_tables(5, 27, pts)[:, [0, 1, 2, 3, 4, 5, 6, 7, 9, 8] + list(range(10, 27))]
    ref = py_oracle.lagrange2(5, pts)
    assert np.abs(got - ref).max() > 1e-3


def membrane_systems(ctx):
    p = sa.Problem(ctx, "membrane3d", DISK)
    m = sa.Problem(ctx, "membrane3d", DISK + ";measure=1")
    assert p.info()["ndofs"] == m.info()["ndofs"]
    nd = p.info()["ndofs"]
    cp = np.ctypeslib.as_array(p.desc().constrained, shape=(nd,))
    cm = np.ctypeslib.as_array(m.desc().constrained, shape=(nd,))
    assert np.array_equal(cp, cm)
    assert all(p.desc().buckets[b].n_geo == 27 for b in range(p.desc().n_buckets))
    return p, m, cp.astype(bool)


def test_membrane_golden_oracle():
    p, m, cons = membrane_systems(None)
    info = p.info()
    nd, nf = info["ndofs"], info["nfree"]
    o = py_oracle.assemble(p.desc_ptr(), p.numeric(), nd, nf)
    A = sp.csr_matrix((o.Aval, o.Acol, o.Arow), shape=(nf, nf))
    x = np.zeros(nd)
    x[~cons] = spl.spsolve(A.tocsc(), o.rhs[~cons])
    om = py_oracle.assemble(m.desc_ptr(), m.numeric(), nd, nf)
    val = float(np.dot(om.rhs, x))
    print("integral of u_z: %.9e" % val)
    assert LO < val < HI


@pytest.mark.gpu
@pytest.mark.parametrize("coef_mode", ["auto", "compiled"])
def test_membrane_golden_gpu(ctx, coef_mode):
    # curved Jacobians on the device (compiled coefficient programs: the
    # program needs more registers than the interpreter holds, so the default
    # mode compiles it too), FP64 assembly against the oracle, device solve,
    # golden window
    print("case membrane gpu coef_mode", coef_mode, flush=True)
    p, m, cons = membrane_systems(ctx)
    info = p.info()
    nd, nf = info["ndofs"], info["nfree"]
    plan = p.plan()
    plan.coef_mode(coef_mode)
    if coef_mode == "auto":
        with pytest.raises(sa.SlxError, match="interpreter holds"):
            plan.coef_mode("interpreter")
    g = p.generate("fp64").csr(nd)
    o = py_oracle.assemble(p.desc_ptr(), p.numeric(), nd, nf)
    assert np.array_equal(g.Acol, o.Acol) and np.array_equal(g.Arow, o.Arow)
    ref = np.sqrt(np.sum(o.Aval ** 2) + np.sum(o.Dval ** 2))
    err = np.sqrt(np.sum((g.Aval - o.Aval) ** 2) + np.sum((g.Dval - o.Dval) ** 2)) / ref
    assert err <= 1e-12 and np.linalg.norm(g.rhs - o.rhs) <= 1e-12 * np.linalg.norm(o.rhs)
    x, it = plan.solve(nd, rel_tol=1e-12, max_outer=30, max_inner=20000)
    assert it["rel_residual"] <= 1e-11
    mplan = m.plan()
    mplan.coef_mode(coef_mode)
    bm = m.generate("fp64").csr(nd).rhs
    val = float(np.dot(bm, x))
    print("integral of u_z (GPU FP64 assembly + device solve): %.9e" % val)
    assert LO < val < HI
    # the FP16 split on the same curved mesh, within its gate
    s = p.generate("f16split").csr(nd)
    es = np.sqrt(np.sum((s.Aval - o.Aval) ** 2) + np.sum((s.Dval - o.Dval) ** 2)) / ref
    print("f16 split global error %.2e" % es)
    assert es <= 1e-6
