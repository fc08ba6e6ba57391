"""Orientation folding against the reference's orientation batches.

The reference integrates every total orientation as its own batch with the
tables of that orientation (elementselector.cpp:4-133, contribution.cpp:61-133).
A folded plan integrates all orientations of a (type, order) group with the
orientation-0 tables, renumbers the functions in the address tables and signs
the rows/columns.  For hexahedra and quadrangles the edge/face functions of any
orientation are signed renumberings of orientation 0 (Legendre parity,
h1hexahedron.cpp:134-205), so both give the same matrix up to rounding.

CPU: the oracle on the folded description equals the oracle on the unfolded
one (one batch per orientation).  GPU: the HIP path on the folded plan equals
the oracle on the unfolded description.
"""
import numpy as np
import pytest

import sparselizard_amd as sa
from oracle import py_oracle

CASES = [
    ("hppoisson3d", "n=3;pmin=1;pmax=4"),
    ("hppoisson3d", "n=2;pmin=5;pmax=5"),
    ("projection", "elem=hex;n=2;order=4"),
    ("poisson2d", "n=4;order=4"),
    ("poisson2d", "n=4;order=3"),
]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def oracle_csr(kind, spec, fold):
    p = sa.Problem(None, kind, spec + ";fold=%d" % fold)
    info = p.info()
    return p, py_oracle.assemble(p.desc_ptr(), p.numeric(), info["ndofs"], info["nfree"])


def check_same(got, ref, tol):
    for name in ("Arow", "Acol", "Drow", "Dcol"):
        assert np.array_equal(getattr(got, name), getattr(ref, name)), name
    for name in ("Aval", "Dval", "rhs"):
        b = getattr(ref, name)
        if b.size:
            e = rel(getattr(got, name), b)
            assert e <= tol, "%s rel err %.3e" % (name, e)


@pytest.mark.parametrize("kind,spec", CASES)
def test_folded_equals_orientation_batches(kind, spec):
    p0, o0 = oracle_csr(kind, spec, 0)
    p1, o1 = oracle_csr(kind, spec, 1)
    assert p1.desc().n_buckets <= p0.desc().n_buckets
    check_same(o1, o0, 1e-13)


def test_hex_orientations_fold_into_order_buckets():
    # one matrix and one rhs bucket per order, whatever the number of orientations
    p0 = sa.Problem(None, "hppoisson3d", "n=4;pmin=2;pmax=4;fold=0")
    p1 = sa.Problem(None, "hppoisson3d", "n=4;pmin=2;pmax=4")
    assert p0.desc().n_buckets > 20
    assert p1.desc().n_buckets <= 2 * 3
    d = p1.desc()
    signed = [d.buckets[b].slices[s].tf_sign for b in range(d.n_buckets) for s in range(d.buckets[b].n_slices)]
    assert any(bool(x) for x in signed)


def test_tet_p2_has_no_orientation_batches():
    # order-2 bases do not depend on orientation: nothing to fold, no signs
    p = sa.Problem(None, "elasticity3d", "n=2")
    d = p.desc()
    for b in range(d.n_buckets):
        for s in range(d.buckets[b].n_slices):
            assert not d.buckets[b].slices[s].tf_sign and not d.buckets[b].slices[s].dof_sign


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("f16split", 1e-6), ("fp32", 2e-5)])
@pytest.mark.parametrize("kind,spec", CASES[:3])
def test_gpu_folded_plan_vs_orientation_batches(ctx, kind, spec, precision, tol):
    print("case fold", kind, spec, precision, flush=True)
    _, ref = oracle_csr(kind, spec, 0)
    p = sa.Problem(ctx, kind, spec)
    info = p.info()
    got = p.generate(precision).csr(info["ndofs"])
    check_same(got, ref, tol)
