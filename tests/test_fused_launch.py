"""One persistent launch over the factorized slices of every bucket
(kernels/factor.hip factor_fused_kernel / factor64_fused_kernel): the element
groups that contribution::generate visits one after the other
(contribution.cpp:61-133) share one work queue, highest per-element cost
first.  Every (slice, element) item runs the same per-element code as the
per-slice launches, so the element matrices and the CSR values must be
bitwise those of the per-slice launches (SLX_TUNE_FUSED=0).  The FP16 and
BF16 splits use the fused launch by default, FP32 MFMA and FP64 with
SLX_TUNE_FUSED=1 (measured slower there); parity against
the oracle is covered by the MFMA tests, which run the fused launch by
default on mixed-order hexahedral meshes."""
import ctypes as C

import numpy as np
import pytest

import sparselizard_amd as sa

pytestmark = pytest.mark.gpu

MODES = ["f16split", "bf16split", "fp32mfma", "fp64"]


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


def assemble(ctx, spec, precision):
    prob = sa.Problem(ctx, "hppoisson3d", spec)
    plan = prob.generate(precision)
    return prob, plan, plan.element_matrices()[entries(prob, plan)], plan.csr(prob.info()["ndofs"])


@pytest.mark.parametrize("precision", MODES)
def test_fused_equals_per_slice_launches(ctx, precision, monkeypatch):
    print("case fused", precision, flush=True)
    spec = "n=3;pmin=1;pmax=5"
    monkeypatch.setenv("SLX_TUNE_FUSED", "1")  # every arithmetic (FP32 MFMA and FP64 default to per-slice launches)
    p1, plan1, ke1, c1 = assemble(ctx, spec, precision)
    n_slices, n_items = plan1.fused_info(precision)
    assert n_slices >= 2 and n_items > 0, "no fused launch on a mixed-order mesh"
    monkeypatch.setenv("SLX_TUNE_FUSED", "0")
    p2, plan2, ke2, c2 = assemble(ctx, spec, precision)
    assert plan2.fused_info(precision) == (0, 0)
    assert np.array_equal(ke1, ke2)
    for name in ("Aval", "Dval", "rhs"):
        assert np.array_equal(getattr(c1, name), getattr(c2, name)), name


def test_fused_graph_replay(ctx):
    # the fused launch (its work counter cleared by a kernel, not a memset)
    # captured in a HIP graph: every replay equals the eager assembly, with
    # eager work between the replays
    import torch
    print("case fused graph", flush=True)
    dev = torch.device("cuda:0")
    p = sa.Problem(ctx, "hppoisson3d", "n=3;pmin=1;pmax=5")
    info = p.info()
    nd = p.numeric()
    plan = p.plan()
    assert plan.fused_info("f16split")[0] >= 2
    xyz = torch.from_numpy(np.ctypeslib.as_array(nd.node_xyz, shape=(3 * p.n_nodes(),)).copy()).to(dev)
    par_n = nd.n_params * nd.param_len
    par = torch.from_numpy(np.ctypeslib.as_array(nd.params, shape=(par_n,)).copy() if par_n else np.zeros(1)).to(dev)
    fptrs = (C.c_void_p * 1)(None)
    torch.cuda.synchronize()
    ref = p.generate("f16split").csr(info["ndofs"])
    plan.graph_capture(C.c_void_p(xyz.data_ptr()), C.c_void_p(par.data_ptr()), nd.param_len, fptrs, "f16split")
    for rep in range(3):
        plan.graph_launch()
        plan.sync()
        got = plan.csr(info["ndofs"])
        assert np.array_equal(got.Aval, ref.Aval) and np.array_equal(got.rhs, ref.rhs), "replay %d differs" % rep
        p.generate("fp64")  # eager work between replays
        ref = p.generate("f16split").csr(info["ndofs"])
