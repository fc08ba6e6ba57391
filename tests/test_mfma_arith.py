"""BF16 hi/mid/lo split MFMA and plain FP32 MFMA element contractions against
the CPU oracle (restated contribution::generate + rawmat::process,
contribution.cpp:172-246, rawmat.cpp:127-298), through the C ABI.

Gates: the CSR pattern bit-exact in every case; global relative Frobenius
error of A, D and the rhs <= 1e-6 for the BF16 split (the same gate as the FP16
split) and <= 2e-5 for FP32 MFMA (the gate of the FP32 VALU element kernel);
element matrices slice by slice within the same gates."""
import numpy as np
import pytest

import sparselizard_amd as sa
from oracle import py_oracle
from _replay import oracle_ke_layout

pytestmark = pytest.mark.gpu

TOL_BF16 = 1e-6
TOL_F32 = 2e-5
MODES = [("bf16split", TOL_BF16), ("fp32mfma", TOL_F32)]


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def check_case(ctx, kind, spec, precision, tol):
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
    print("errors", errs, flush=True)
    for name, e in errs.items():
        assert e <= tol, "%s rel err %.3e > %.1e" % (name, e, tol)
    return prob, plan, got, errs


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_operand_layout(ctx, kind):
    # bf16 16x16x32, f32 16x16x4, f64 16x16x4 on small integer operands:
    # D = A B exactly under the lane layouts the kernels assume
    from sparselizard_amd import _lib as L
    d = L.mfma_layout_probe(ctx, kind)
    kk = 32 if kind == 1 else 4
    m = np.arange(16)[:, None]
    k = np.arange(kk)[None, :]
    A = ((m + 2 * k) % 7 - 3).astype(np.float64)
    k2 = np.arange(kk)[:, None]
    n = np.arange(16)[None, :]
    B = ((3 * k2 + n) % 5 - 2).astype(np.float64)
    assert np.array_equal(d.astype(np.float64), A @ B)


@pytest.mark.parametrize("precision,tol", MODES)
@pytest.mark.parametrize("kind,spec", [
    ("elasticity3d", "n=3"),                  # tet P2 h1xyz: batched kernel, 9 slices of 10 x 10
    ("elasticity3d", "n=4;order=1"),          # tet P1 h1xyz: 4 x 4 slices
    ("hppoisson3d", "n=3"),                   # orders 1..5: batched (p = 1) and factorized (p >= 2)
    ("hppoisson3d", "n=2;pmin=5;pmax=5"),     # p = 5 only, folded orientations (signed slots)
])
def test_mfma_modes_match_oracle(ctx, kind, spec, precision, tol, monkeypatch):
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    prob, plan, got, errs = check_case(ctx, kind, spec, precision, tol)
    kc = plan.kernel_counts(precision)
    assert kc["batched"] + kc["factorized"] >= 1, "no MFMA kernel in %s" % precision


@pytest.mark.parametrize("precision,tol", MODES)
def test_mfma_modes_triangles(ctx, precision, tol, monkeypatch):
    # 2D P1 triangles with the table off: the batched kernel (the tuned table
    # sends this class to the VALU kernel)
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    prob, plan, got, errs = check_case(ctx, "poisson2d", "n=9", precision, tol)
    assert plan.kernel_counts(precision)["batched"] >= 1


@pytest.mark.parametrize("precision,tol", MODES)
def test_mfma_modes_qvarying_on_valu(ctx, precision, tol):
    # Gauss-point-varying coefficients: the FP32 VALU element kernel in these modes
    check_case(ctx, "quadbatch", "instances=3;nmin=2;nmax=5", precision, tol)


@pytest.mark.parametrize("precision,tol", MODES)
def test_mfma_modes_prism_factorized(ctx, precision, tol):
    # order-5 prisms (126 x 126 element matrices, Gauss-point-varying
    # Jacobians): the factorized kernel with coefficients per Gauss point
    prob, plan, got, errs = check_case(ctx, "prism3d", "n=2;order=5", precision, tol)


@pytest.mark.parametrize("precision,tol", MODES)
@pytest.mark.parametrize("kind,spec", [("elasticity3d", "n=2"), ("hppoisson3d", "n=2"), ("poisson2d", "n=5")])
def test_mfma_modes_element_matrices(ctx, kind, spec, precision, tol, monkeypatch):
    """Element matrices slice by slice against the oracle's."""
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    print("case ke", kind, spec, precision, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    plan = prob.generate(precision)
    ke = plan.element_matrices()
    kb, kl = plan.ke_layout()
    o, ke_ref = oracle_ke_layout(prob, kb, kl, len(ke))
    d = prob.desc()
    worst, k = [], 0
    for b in range(d.n_buckets):
        bd = d.buckets[b]
        for s in range(bd.n_slices):
            sd = bd.slices[s]
            ndf = 1 if sd.is_rhs else sd.n_dof
            e, i, j = np.meshgrid(np.arange(bd.n_elem), np.arange(sd.n_tf), np.arange(ndf), indexing="ij")
            idx = (kb[k] + (e * sd.n_tf + i) * kl[k] + j).ravel()
            worst.append(relerr(ke[idx], ke_ref[idx]))
            k += 1
    print("per-slice element-matrix errors", ["%.2e" % w for w in worst])
    assert max(worst) <= tol


@pytest.mark.parametrize("precision", ["bf16split", "fp32mfma"])
def test_mfma_modes_deterministic(ctx, precision):
    prob = sa.Problem(ctx, "elasticity3d", "n=3")
    info = prob.info()
    a = prob.generate(precision).csr(info["ndofs"])
    b = prob.generate(precision).csr(info["ndofs"])
    assert np.array_equal(a.Aval, b.Aval) and np.array_equal(a.Dval, b.Dval) and np.array_equal(a.rhs, b.rhs)


@pytest.mark.parametrize("kind,spec", [
    ("hppoisson3d", "n=3"),                   # orders 1..5: factorized FP64 for p >= 2
    ("hppoisson3d", "n=2;pmin=5;pmax=5"),     # p = 5 only, folded orientations
    ("prism3d", "n=2;order=5"),               # Gauss-point-varying coefficients
])
def test_fp64_factorized_matches_oracle(ctx, kind, spec):
    # the factorized contraction on v_mfma_f64_16x16x4_f64: FP64 gate
    prob, plan, got, errs = check_case(ctx, kind, spec, "fp64", 1e-12)
    assert plan.kernel_counts("fp64")["factorized"] >= 1


def test_fp64_factorized_element_matrices_and_valu(ctx, monkeypatch):
    """FP64 factorized element matrices slice by slice against the oracle's
    (1e-12), bitwise repeatable, and equal to the FP64 VALU kernel's to
    rounding (a different summation order)."""
    print("case ke fp64 factorized", flush=True)
    prob = sa.Problem(ctx, "hppoisson3d", "n=2")
    plan = prob.generate("fp64")
    assert plan.kernel_counts("fp64")["factorized"] >= 1
    ke = plan.element_matrices()
    kb, kl = plan.ke_layout()
    o, ke_ref = oracle_ke_layout(prob, kb, kl, len(ke))
    d = prob.desc()
    idx_all, k = [], 0
    for b in range(d.n_buckets):
        bd = d.buckets[b]
        for s in range(bd.n_slices):
            sd = bd.slices[s]
            ndf = 1 if sd.is_rhs else sd.n_dof
            e, i, j = np.meshgrid(np.arange(bd.n_elem), np.arange(sd.n_tf), np.arange(ndf), indexing="ij")
            idx = (kb[k] + (e * sd.n_tf + i) * kl[k] + j).ravel()
            assert relerr(ke[idx], ke_ref[idx]) <= 1e-12
            idx_all.append(idx)
            k += 1
    idx = np.concatenate(idx_all)
    again = prob.generate("fp64").element_matrices()
    assert np.array_equal(ke[idx], again[idx])
    monkeypatch.setenv("SLX_TUNE_FP64_FACTOR", "0")
    prob2 = sa.Problem(ctx, "hppoisson3d", "n=2")
    plan2 = prob2.generate("fp64")
    assert plan2.kernel_counts("fp64")["factorized"] == 0
    assert relerr(ke[idx], plan2.element_matrices()[idx]) <= 1e-13


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("precision,tol", [("f16split", 1e-6), ("fp64", 1e-12)])
def test_factor_tile_shapes(ctx, variant, precision, tol, tmp_path, monkeypatch):
    # every tile shape of the factorized kernel the kernel table can name for
    # order-3 hexahedra (64 x 64 element matrices): (2, 1), (4, 2), (7, 2)
    # tiles per wave, same results to the mode's gate
    table = tmp_path / "table.txt"
    table.write_text("5 64 64 0 factor 1 %d\n" % variant)
    monkeypatch.setenv("SLX_TUNE_TABLE", str(table))
    prob, plan, got, errs = check_case(ctx, "hppoisson3d", "n=2;pmin=3;pmax=3", precision, tol)
    assert plan.kernel_counts(precision)["factorized"] >= 1
