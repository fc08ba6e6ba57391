"""GPU parity: the HIP assembly path against the CPU FP64 oracle.

Pattern (rowptr / column indices of A and D) must be bit-exact; values match
to the stated relative Frobenius tolerance.  Every case runs through the C ABI
(slx_problem_generate -> slx_plan_create / slx_assemble / slx_csr_copyout).
"""
import ctypes as C

import numpy as np
import pytest

import sparselizard_amd as sa
from oracle import py_oracle
from _replay import oracle_ke_layout, reduce_host

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
def test_f16split_mfma_matches_oracle(ctx, kind, spec, monkeypatch):
    # quadbatch has coefficients that vary over the Gauss points: by default its
    # buckets go to the FP32 VALU kernel; the knob keeps them on the batched
    # MFMA kernel so its per-Gauss-point column path stays covered
    # (and the tuned kernel table, which may prefer the VALU kernel for some classes, is off)
    monkeypatch.setenv("SLX_TUNE_MFMA_QVARYING", "1")
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    prob, plan, got, errs = run_case(ctx, kind, spec, "f16split", TOL_SPLIT)
    assert plan.kernels()["mfma_buckets"] >= 1, "MFMA kernel not used"
    print("errors", errs)


def test_f16split_qvarying_buckets_on_valu(ctx, monkeypatch):
    # with the one-tile kernel off (SLX_TUNE_TILE=0), Gauss-point-varying
    # coefficients run on the FP32 VALU element kernel, still within the split
    # path's 1e-6 gate (the one-tile kernel has its own tests)
    monkeypatch.setenv("SLX_TUNE_TILE", "0")
    prob, plan, got, errs = run_case(ctx, "quadbatch", "instances=3;nmin=2;nmax=5", "f16split", TOL_SPLIT)
    assert plan.kernels()["mfma_buckets"] == 0
    assert plan.tile_count("f16split") == 0
    print("errors", errs)


@pytest.mark.parametrize("spec", ["n=5;jitter=0.15", "n=3;order=2", "n=3;order=3", "n=2;order=5"])
@pytest.mark.parametrize("precision,tol", [("fp64", TOL_FP64), ("fp32", TOL_FP32), ("f16split", TOL_SPLIT)])
def test_prism3d(ctx, spec, precision, tol):
    # prisms: Gauss-point-varying Jacobians, stiffness + mass + rhs; order 3 has
    # orientation-dependent face functions, order 5 (126 x 126 element matrices)
    # takes the factorized kernel in the split mode
    run_case(ctx, "prism3d", spec, precision, tol)


def test_prism3d_f16split_mfma(ctx, monkeypatch):
    monkeypatch.setenv("SLX_TUNE_MFMA_QVARYING", "1")
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    prob, plan, got, errs = run_case(ctx, "prism3d", "n=4;jitter=0.15", "f16split", TOL_SPLIT)
    assert plan.kernels()["mfma_buckets"] >= 1, "MFMA kernel not used"


@pytest.mark.parametrize("precision", ["fp64", "fp32", "f16split"])
def test_every_dof_constrained(ctx, precision):
    # one square, two triangles, all four nodes on the Dirichlet boundary: no
    # free DOF, empty A and D; the element kernels still run, the reduction has
    # no rows, and the solve returns the Dirichlet values
    prob, plan, got, errs = run_case(ctx, "poisson2d", "n=1", precision, TOL_FP64)
    info = prob.info()
    assert info["nfree"] == 0 and len(got.Aval) == 0 and len(got.Dval) == 0
    assert np.array_equal(got.Arow, [0]) and np.array_equal(got.Drow, [0])
    x, si = plan.solve(info["ndofs"])
    assert np.array_equal(x, got.rhs) and si["inner"] == 0
    y = plan.matvec(np.arange(info["ndofs"], dtype=np.float64))
    assert np.array_equal(y, np.arange(info["ndofs"], dtype=np.float64))


@pytest.mark.parametrize("precision,tol", [("fp64", TOL_FP64), ("fp32", TOL_FP32)])
def test_hp_hex_long_segments(ctx, precision, tol):
    # per-element orders 1..5 on hexahedra: element rows up to 216 entries
    # (strided reduction) and rows longer than 255 columns (16-bit slots)
    run_case(ctx, "hppoisson3d", "n=3", precision, tol)


@pytest.mark.parametrize("elem", ["tri", "quad", "tet", "hex"])
def test_projection_fp32_qvarying(ctx, elem):
    # non-affine quads/hexes vary over Gauss points; FP32 VALU element kernel
    run_case(ctx, "projection", "elem=%s;n=3;order=2" % elem, "fp32", TOL_FP32)


@pytest.mark.parametrize("precision", ["fp64", "f16split"])
def test_deterministic_repeat(ctx, precision):
    prob = sa.Problem(ctx, "elasticity3d", "n=3")
    info = prob.info()
    a = prob.generate(precision).csr(info["ndofs"])
    b = prob.generate(precision).csr(info["ndofs"])
    assert np.array_equal(a.Aval, b.Aval) and np.array_equal(a.Dval, b.Dval) and np.array_equal(a.rhs, b.rhs)


@pytest.mark.parametrize("kind,spec", [("poisson2d", "n=5"), ("elasticity3d", "n=2"), ("hppoisson3d", "n=2"),
                                       ("prism3d", "n=3"), ("prism3d", "n=2;order=3")])
@pytest.mark.parametrize("precision,tol", [("fp64", TOL_FP64), ("f16split", TOL_SPLIT), ("fp32", TOL_FP32)])
def test_element_matrices_match_oracle(ctx, kind, spec, precision, tol):
    """Element matrices (before the reduction) against the oracle, slice by
    slice, and the device reduction against its host replay on the device's
    own element matrices."""
    print("case ke", kind, spec, precision, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    info = prob.info()
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
    # the device reduction applied to the device's element matrices
    host = sa.Problem(None, kind, spec)
    hplan = host.plan()
    maps = hplan.reduction_maps()
    assert np.array_equal(maps["ke_base"], kb) and np.array_equal(maps["ke_ld"], kl)
    Aval, Dval, rhs_free = reduce_host(maps, ke, o.Arow, o.Drow, hplan.shape())
    got = plan.csr(info["ndofs"])
    red_err = relerr(got.Aval, Aval)
    print("device reduction vs host replay on device Ke: %.2e" % red_err)
    assert max(worst) <= tol, "slice errors %s" % worst
    assert red_err <= 1e-13


@pytest.mark.parametrize("kind,spec", [("poisson2d", "n=5"), ("elasticity3d", "n=2"), ("hppoisson3d", "n=2"),
                                       ("projection", "elem=hex;n=2;order=2"), ("prism3d", "n=3")])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
def test_coefficient_kernel_matches_oracle(ctx, kind, spec, precision, tol):
    """Gauss-point coefficients (geometry + program) of the last bucket."""
    print("case coef", kind, spec, precision, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    plan = prob.generate(precision)
    got = plan.coefficients()
    d = prob.desc()
    ref = py_oracle.coefficients(prob.desc_ptr(), prob.numeric(), d.n_buckets - 1, len(got))
    bd = d.buckets[d.n_buckets - 1]
    nq = 1 if bd.q_invariant else bd.n_gp
    per_slot = [relerr(got[c * nq * bd.n_elem:(c + 1) * nq * bd.n_elem], ref[c * nq * bd.n_elem:(c + 1) * nq * bd.n_elem])
                for c in range(bd.n_coefs)]
    print("per-slot coefficient errors", ["%.1e" % x for x in per_slot])
    assert max(per_slot) <= tol


def test_plan_keeps_its_context_alive():
    # slx_destroy on a context with live plans only drops its reference: the
    # plans keep assembling on the context's stream until they are destroyed
    print("case context lifetime", flush=True)
    c = sa.Context(0)
    prob = sa.Problem(c, "poisson2d", "n=6")
    info = prob.info()
    prob.generate("fp64")
    first = prob.plan().csr(info["ndofs"])
    c.close()
    again = prob.generate("fp64").csr(info["ndofs"])
    assert np.array_equal(first.Aval, again.Aval)
    prob.close()


def test_assemble_async_on_caller_stream(ctx):
    # slx_assemble_async on a torch stream gives the matrix of slx_assemble
    import torch
    print("case assemble_async", flush=True)
    prob = sa.Problem(ctx, "elasticity3d", "n=3")
    info = prob.info()
    ref = prob.generate("f16split").csr(info["ndofs"])
    plan = prob.plan()
    s = torch.cuda.Stream()
    plan.assemble_async(prob.numeric(), "f16split", C.c_void_p(s.cuda_stream))
    s.synchronize()
    got = plan.csr(info["ndofs"])
    assert np.array_equal(got.Aval, ref.Aval) and np.array_equal(got.rhs, ref.rhs)
