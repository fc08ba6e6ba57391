"""One wave per element holding its 16x16 MFMA tiles (kernels/tile.hip) for the
small matrix slices whose coefficients vary over the Gauss points, in the FP16
split, the BF16 split and FP32 MFMA: the
element-matrix contraction of contribution.cpp:172-246 regrouped by tf
derivative, Ke = P^T Q_e with the shared tf table P and the element's
Q_e[k][j] = w_q sum_b D_e[q][a][b] dof_b[j][q].

Against the CPU FP64 oracle: the CSR pattern bit-exact and the values within
the split path's 1e-6 relative Frobenius gate, on every element family that
takes the kernel (order-3 quadrangles, P2 triangles with a Newton field, order-1
prisms and hexahedra), plus the element matrices against the FP32 VALU kernel
(SLX_TUNE_TILE=0) that handles the same slices otherwise."""
import numpy as np
import pytest

import sparselizard_amd as sa
from oracle import py_oracle

pytestmark = pytest.mark.gpu

TOL_SPLIT = 1e-6


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def entries(prob, plan):
    """Indices of every element-matrix entry (padding between segments excluded)."""
    kb, kl = plan.ke_layout()
    d = prob.desc()
    idx, k = [], 0
    for b in range(d.n_buckets):
        bd = d.buckets[b]
        for s in range(bd.n_slices):
            sd = bd.slices[s]
            ndf = 1 if sd.is_rhs else sd.n_dof
            e, i, j = np.meshgrid(np.arange(bd.n_elem), np.arange(sd.n_tf), np.arange(ndf), indexing="ij")
            idx.append((kb[k] + (e * sd.n_tf + i) * kl[k] + j).ravel())
            k += 1
    return np.concatenate(idx)


# one tile (quadrangles order 3, P2 triangles, order-1 prisms and hexahedra),
# 2 x 2 tiles (order-2 hexahedra 27 x 27, order-2 prisms 18 x 18, order-2 hcurl
# tetrahedra 30 x 30 with value and curl tables), 3 x 3 (order-3 prisms 40 x 40
# at 80 Gauss points: two passes of the lanes over the points) and 4 x 4
# (order-3 hexahedra 64 x 64)
CASES = [("quadbatch", "instances=3;nmin=2;nmax=5"), ("newton2d", "n=6"), ("prism3d", "n=3"),
         ("hppoisson3d", "n=3;pmin=1;pmax=2"), ("quadbatch", "instances=2;nmin=7;nmax=7"),
         ("prism3d", "n=2;order=2"), ("hcurlprojection", "elem=tet;n=2;order=2;term=maxwell"),
         ("hppoisson3d", "n=2;pmin=2;pmax=2"), ("prism3d", "n=2;order=3"), ("hppoisson3d", "n=2;pmin=3;pmax=3")]


@pytest.mark.parametrize("precision", ["f16split", "bf16split", "fp32mfma"])
@pytest.mark.parametrize("kind,spec", CASES)
def test_one_tile_kernel_matches_oracle(ctx, kind, spec, precision, monkeypatch):
    # every slice size (by default the kernel takes n_tf * n_dof >= 100 only),
    # whatever the kernel table names for the class; the three MFMA arithmetics
    # (gates: 1e-6 for the splits, 2e-5 for FP32 as for the FP32 VALU kernel)
    monkeypatch.setenv("SLX_TUNE_TILE_MIN", "0")
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    print("case tile", kind, spec, precision, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    info = prob.info()
    plan = prob.generate(precision)
    assert plan.tile_count(precision) >= 1, "one-tile kernel not used"
    assert plan.tile_count("fp32") == 0 and plan.tile_count("fp64") == 0
    tol = 2e-5 if precision == "fp32mfma" else TOL_SPLIT
    got = plan.csr(info["ndofs"])
    ref = py_oracle.assemble(prob.desc_ptr(), prob.numeric(), info["ndofs"], info["nfree"])
    for name in ("Arow", "Acol", "Drow", "Dcol"):
        assert np.array_equal(getattr(got, name), getattr(ref, name)), name
    for name in ("Aval", "Dval", "rhs"):
        b = getattr(ref, name)
        if b.size:
            e = relerr(getattr(got, name), b)
            print(name, "rel err %.3e" % e)
            assert e <= tol, "%s rel err %.3e" % (name, e)


@pytest.mark.parametrize("kind,spec", CASES[:3] + CASES[5:7] + CASES[8:])
def test_one_tile_element_matrices_match_valu(ctx, kind, spec, monkeypatch):
    # same slices, same layout: the one-tile kernel's element matrices against
    # those of the kernel that takes the slices without it (the FP32 VALU kernel,
    # or the factorized MFMA kernel for 2 x 2 tiles), both against the FP64
    # path's (1e-6 of its norm)
    print("case tile vs valu", kind, spec, flush=True)
    monkeypatch.setenv("SLX_TUNE_TILE_MIN", "0")
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    prob = sa.Problem(ctx, kind, spec)
    plan = prob.generate("fp64")
    sel = entries(prob, plan)
    k64 = plan.element_matrices()[sel]
    ktile = prob.generate("f16split").element_matrices()[sel]
    monkeypatch.setenv("SLX_TUNE_TILE", "0")
    prob2 = sa.Problem(ctx, kind, spec)
    plan2 = prob2.generate("f16split")
    assert plan2.tile_count("f16split") == 0
    kvalu = plan2.element_matrices()[entries(prob2, plan2)]
    assert ktile.shape == kvalu.shape == k64.shape
    e_tile, e_valu = relerr(ktile, k64), relerr(kvalu, k64)
    print("element matrices: tile %.3e, other kernel %.3e" % (e_tile, e_valu))
    assert e_tile <= TOL_SPLIT and e_valu <= TOL_SPLIT


def test_one_tile_default_slice_sizes(ctx, monkeypatch):
    # default threshold: order-3 quadrangles (16 x 16) take the kernel, P2
    # triangles (6 x 6) stay on the VALU element kernel
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    assert sa.Problem(ctx, "quadbatch", "instances=2;nmin=2;nmax=3").generate("f16split").tile_count("f16split") >= 1
    assert sa.Problem(ctx, "newton2d", "n=4").generate("f16split").tile_count("f16split") == 0


def test_kernel_table_decides(ctx, tmp_path, monkeypatch):
    # a table entry naming another kernel keeps the class off the one-tile kernel;
    # a "tile" entry or none leaves it there
    # (the library reads a table again when SLX_TUNE_TABLE names another file)
    valu, tile = tmp_path / "valu.txt", tmp_path / "tile.txt"
    valu.write_text("3 16 16 0 valu 1 -1\n")
    tile.write_text("3 16 16 0 tile 1 -1\n")
    monkeypatch.setenv("SLX_TUNE_TABLE", str(valu))
    assert sa.Problem(ctx, "quadbatch", "instances=2;nmin=2;nmax=3").generate("f16split").tile_count("f16split") == 0
    monkeypatch.setenv("SLX_TUNE_TABLE", str(tile))
    assert sa.Problem(ctx, "quadbatch", "instances=2;nmin=2;nmax=3").generate("f16split").tile_count("f16split") >= 1
