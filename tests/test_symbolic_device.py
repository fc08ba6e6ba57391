"""The device symbolic phase (kernels/symbolic.hip) against the host phase
(plan.cpp build_symbolic, the restated rawmat::process of rawmat.cpp:127-298
with the A/D split of gentools.cpp:2040-2068) on the same description: the CSR
pattern and every reduction array bit for bit, and the assembled values
bitwise equal.  The host phase itself is checked against the oracle's
restated rawmat::process in the CPU suite (test_cpu_reference.py) and through
every GPU parity test that ran before the device phase existed.

Cases cover one and several fields, constraints on and off, every DOF
constrained (no free row), folded orientations (signed slots), per-element
orders 1..5 on hexahedra (rows of up to 1728 entries: the middle tier of the
column kernel), order-4 vector tetrahedra (rows beyond 2048 entries: the
large tier), prisms, rhs-only and Newton-field problems, and the two
reference example meshes."""
import os

import numpy as np
import pytest

import sparselizard_amd as sa
from sparselizard_amd._lib import Plan, PLAN_HOST_SYMBOLIC, SYMBOLIC_ARRAYS

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [
    ("poisson2d", "n=1"),
    ("poisson2d", "n=40"),
    ("poisson2d", "n=12;order=3"),
    ("elasticity3d", "n=4"),
    ("elasticity3d", "n=2;order=1;clamp=0"),
    ("elasticity3d", "n=2;order=4"),
    ("hppoisson3d", "n=4;pmin=2;pmax=4"),
    ("hppoisson3d", "n=4;pmin=2;pmax=4;fold=0"),
    ("hppoisson3d", "n=3;pmin=5;pmax=5"),
    ("quadbatch", "instances=6"),
    ("newton2d", "n=10"),
    ("prism3d", "n=3;order=3"),
    ("convdiff2d", "n=12;order=2;bx=20;by=10"),
    ("projection", "elem=tet;n=3;order=3"),
    ("hcurlprojection", "elem=tet;n=3;order=2;term=maxwell"),
    ("hcurlprojection", "elem=hex;n=3;order=1;term=maxwell"),
    ("electrostatics1d", "file=" + os.path.join(GOLDEN, "electrostatics1d_line.msh")),
    ("hoprojection2d", "file=%s;order=5" % os.path.join(GOLDEN, "hointerp2d_quad.msh")),
]


def csr_pattern(plan, nd):
    c = plan.csr(nd)
    return c.Arow, c.Acol, c.Drow, c.Dcol


@pytest.mark.parametrize("kind,spec", CASES)
def test_device_symbolic_matches_host(ctx, kind, spec):
    print("case symbolic", kind, spec, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    nd = prob.info()["ndofs"]
    dev = Plan.create(ctx, prob)
    host = Plan.create(ctx, prob, PLAN_HOST_SYMBOLIC)
    assert dev.symbolic_info()[0] and not host.symbolic_info()[0]
    assert dev.shape() == host.shape()
    assert dev.stats() == host.stats()
    for a, b, name in zip(csr_pattern(dev, nd), csr_pattern(host, nd), ("Arow", "Acol", "Drow", "Dcol")):
        assert np.array_equal(a, b), (kind, spec, name)
    for name in SYMBOLIC_ARRAYS:
        a, b = dev.symbolic_array(name), host.symbolic_array(name)
        assert a.size == b.size and np.array_equal(a, b), (kind, spec, name, a.size, b.size)
    num = prob.numeric()
    for precision in ("fp64", "f16split"):
        dev.assemble(num, precision)
        host.assemble(num, precision)
        x, y = dev.csr(nd), host.csr(nd)
        for name in ("Aval", "Dval", "rhs"):
            assert np.array_equal(getattr(x, name), getattr(y, name)), (kind, spec, precision, name)
    dev.close()
    host.close()
    prob.close()


def test_device_symbolic_rejects_bad_addresses(ctx):
    # a dof address beyond n_dofs is refused by the device phase with the host
    # phase's message (planted in the description's address table, restored after)
    prob = sa.Problem(ctx, "poisson2d", "n=4")
    d = prob.desc()
    b, sl = next((bk, bk.slices[s]) for bk in (d.buckets[i] for i in range(d.n_buckets))
                 for s in range(bk.n_slices) if not bk.slices[s].is_rhs)
    dof = np.ctypeslib.as_array(sl.dof_addr, shape=(sl.n_dof * b.n_elem,))
    keep = dof[3]
    dof[3] = d.n_dofs + 5
    try:
        with pytest.raises(sa.SlxError, match="address out of range"):
            Plan.create(ctx, prob)
        with pytest.raises(sa.SlxError, match="address out of range"):
            Plan.create(ctx, prob, PLAN_HOST_SYMBOLIC)
    finally:
        dof[3] = keep
    prob.close()
