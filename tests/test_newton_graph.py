"""Nonlinear re-assembly (Newton) and its HIP-graph replay.

Workload: P2 field on jittered triangles, reluctivity
nu(|grad a|) = 1 + 500 / (1 + |grad a|^2) evaluated at the Gauss points from
the current field data, Jacobian + residual linearized as in the reference's
saturation example (examples/magnetostatics-vector-potential-saturation-2d/
main.cpp:93-94 and its Newton loop :102-127).  Every step re-assembles at the
field data set from the previous solution (sl::setdata(vec)).

CPU: the oracle's Newton iteration converges quadratically.
GPU: per step, FP64 against the oracle and the f16 split against FP64; the
captured graph replayed with new field data against eager assembly (bitwise),
with eager work between the replays.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import sparselizard_amd as sa
from oracle import py_oracle

SPEC = "n=16;f=500"


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def mask_of(prob):
    n = prob.info()["ndofs"]
    d = prob.desc()
    return np.ctypeslib.as_array(d.constrained, shape=(n,)).astype(bool)


def newton_update(csr, x, cons):
    nf = int((~cons).sum())
    A = sp.csr_matrix((csr.Aval, csr.Acol, csr.Arow), shape=(nf, nf))
    b = np.asarray(csr.rhs)[~cons]
    res = np.linalg.norm(b - A @ x[~cons]) / np.linalg.norm(b)
    xn = x.copy()
    xn[~cons] = spl.spsolve(A.tocsc(), b)
    xn[cons] = np.asarray(csr.rhs)[cons]
    return xn, res


def test_oracle_newton_converges_quadratically():
    p = sa.Problem(None, "newton2d", SPEC)
    info = p.info()
    cons = mask_of(p)
    x = np.zeros(info["ndofs"])
    res = []
    for _ in range(6):
        o = py_oracle.assemble(p.desc_ptr(), p.numeric(), info["ndofs"], info["nfree"])
        x, r = newton_update(o, x, cons)
        p.setdata_vec(x)
        res.append(r)
    assert res[1] < 0.2 and res[2] < 1e-2 and res[3] < 1e-4 and res[5] < 1e-12, res


@pytest.mark.gpu
def test_newton_steps_fp64_and_split(ctx, monkeypatch):
    print("case newton2d per-step", flush=True)
    # Gauss-point-varying coefficients: the knob keeps the split path on the
    # batched MFMA kernel (the default runs them on the FP32 VALU kernel)
    monkeypatch.setenv("SLX_TUNE_MFMA_QVARYING", "1")
    monkeypatch.setenv("SLX_TUNE_TABLE", "none")
    p = sa.Problem(ctx, "newton2d", SPEC)
    info = p.info()
    cons = mask_of(p)
    x = np.zeros(info["ndofs"])
    for step in range(5):
        o = py_oracle.assemble(p.desc_ptr(), p.numeric(), info["ndofs"], info["nfree"])
        g64 = p.generate("fp64").csr(info["ndofs"])
        assert np.array_equal(g64.Acol, o.Acol) and np.array_equal(g64.Arow, o.Arow)
        assert relerr(g64.Aval, o.Aval) < 1e-12 and relerr(g64.rhs, o.rhs) < 1e-12
        plan = p.generate("f16split")
        assert plan.kernels()["mfma_buckets"] >= 1
        gmx = plan.csr(info["ndofs"])
        e_a, e_b = relerr(gmx.Aval, g64.Aval), relerr(gmx.rhs, g64.rhs)
        print("step", step, "split vs fp64: A %.2e rhs %.2e" % (e_a, e_b))
        assert e_a < 1e-6 and e_b < 1e-6
        x, r = newton_update(g64, x, cons)
        p.setdata_vec(x)


@pytest.mark.gpu
def test_newton_mixed_jacobian_reaches_fp64_solution(ctx):
    """Newton on the f16-split assembly converges to the FP64 trajectory's
    solution (inexact Jacobian, residual at the assembly's accuracy)."""
    print("case newton2d mixed trajectory", flush=True)
    sols = {}
    for prec in ("fp64", "f16split"):
        p = sa.Problem(ctx, "newton2d", SPEC)
        info = p.info()
        cons = mask_of(p)
        x = np.zeros(info["ndofs"])
        for _ in range(6):
            x, r = newton_update(p.generate(prec).csr(info["ndofs"]), x, cons)
            p.setdata_vec(x)
        sols[prec] = x
    err = relerr(sols["f16split"], sols["fp64"])
    print("solution rel err split vs fp64: %.2e" % err)
    assert err < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "f16split"])
def test_newton_graph_replay_matches_eager(ctx, precision):
    import ctypes as C
    import torch
    print("case newton2d graph", precision, flush=True)
    dev = torch.device("cuda:0")
    p = sa.Problem(ctx, "newton2d", SPEC)
    info = p.info()
    cons = mask_of(p)
    nd = p.numeric()
    plan = p.plan()
    xyz = torch.from_numpy(np.ctypeslib.as_array(nd.node_xyz, shape=(3 * p.n_nodes(),)).copy()).to(dev)
    par_n = nd.n_params * nd.param_len
    par = torch.from_numpy(np.ctypeslib.as_array(nd.params, shape=(par_n,)).copy() if par_n else np.zeros(1)).to(dev)
    fdata = [torch.from_numpy(f.copy()).to(dev) for f in p.field_data()]
    fptrs = (C.c_void_p * len(fdata))(*[f.data_ptr() for f in fdata])
    torch.cuda.synchronize()
    plan.graph_capture(C.c_void_p(xyz.data_ptr()), C.c_void_p(par.data_ptr()), nd.param_len, fptrs, precision)
    x = np.zeros(info["ndofs"])
    for step in range(4):
        # new field data into the captured buffers, then replay
        for k, f in enumerate(p.field_data()):
            fdata[k].copy_(torch.from_numpy(f.copy()))
        torch.cuda.synchronize()
        plan.graph_launch()
        plan.sync()
        got = plan.csr(info["ndofs"])
        # eager assembly of the same state on the same plan (host inputs)
        ref = p.generate(precision).csr(info["ndofs"])
        assert np.array_equal(got.Aval, ref.Aval) and np.array_equal(got.rhs, ref.rhs), "replay %d differs" % step
        x, r = newton_update(ref, x, cons)
        p.setdata_vec(x)
