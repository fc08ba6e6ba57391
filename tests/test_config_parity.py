"""Per-configuration parity of every precision mode against the CPU oracle
(the restated FP64 generate()+A()+b() of contribution.cpp:133-306 and
rawmat.cpp:127-298), at the largest sizes the threaded oracle assembles in a
few seconds, for the five benchmark workloads:

  poisson2d    2D Poisson P1, 200 x 200 x 2 jittered triangles (full size)
  elasticity3d 3D elasticity h1xyz P2, 20^3 x 6 jittered tetrahedra
  hppoisson3d  3D hp-Poisson, 12^3 jittered hexahedra, orders 1..5
  quadbatch    48 independent K + k^2 M problems, order-3 quadrangles
  newton2d     2D Newton re-assembly P2, 150 x 150 x 2 triangles, 4 steps

Reported per mode (the sl::solve-level checks of sl.cpp:1565-1627):
  - CSR pattern bit-exact against the oracle;
  - element matrices: relative Frobenius error over every slice;
  - global matrix (A and D together) and rhs: relative Frobenius error;
  - solution: the device solve (slx_plan_solve, FP64 refinement to 1e-12)
    of the mode's assembly against scipy's sparse LU of the oracle's system,
    relative 2-norm (on a smaller mesh of the same workload);
  - quadbatch: the largest per-instance relative error (instances are the
    connected components of A's graph);
  - newton2d: the Jacobian error at every Newton step, the iterates driven by
    the oracle's FP64 Newton solutions.
Gates: the matrix errors <= 1e-12 (FP64), 1e-6 (FP16 and BF16 splits), 2e-5
(FP32 MFMA and FP32 VALU).  SLX_PARITY_OUT=<file> writes the report as JSON.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import scipy.sparse.linalg as spl

import sparselizard_amd as sa
from oracle import py_oracle
from _replay import oracle_ke_layout

pytestmark = pytest.mark.gpu

MODES = {"fp64": 1e-12, "f16split": 1e-6, "bf16split": 1e-6, "fp32mfma": 2e-5, "fp32": 2e-5}
CONFIGS = {
    "poisson2d": ("n=200", "n=100"),
    "elasticity3d": ("n=20;order=2", "n=8;order=2"),
    "hppoisson3d": ("n=12", "n=6"),
    "quadbatch": ("instances=48", "instances=6"),
}
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def relerr(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def report(config, payload):
    out = os.environ.get("SLX_PARITY_OUT")
    if not out:
        return
    data = {}
    if os.path.exists(out):
        with open(out) as f:
            data = json.load(f)
    data[config] = payload
    with open(out, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def oracle_system(prob):
    info = prob.info()
    o = py_oracle.assemble(prob.desc_ptr(), prob.numeric(), info["ndofs"], info["nfree"], nthreads=THREADS)
    return info, o


def csr_mats(c, nf, nd):
    A = sp.csr_matrix((c.Aval, c.Acol, c.Arow), shape=(nf, nf))
    D = sp.csr_matrix((c.Dval, c.Dcol, c.Drow), shape=(nf, nd - nf))
    return A, D


def matrix_errors(ctx, kind, spec):
    print("case parity", kind, spec, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    info = prob.info()
    nd, nf = info["ndofs"], info["nfree"]
    plan = prob.plan()
    kb, kl = plan.ke_layout()
    o, ke_ref = oracle_ke_layout(prob, kb, kl, plan.stats()["n_ke"])
    # every element-matrix entry of every (bucket, slice), padding excluded
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
    idx = np.concatenate(idx)
    ke_ref = ke_ref[idx]
    ref_glob = np.sqrt(np.sum(o.Aval ** 2) + np.sum(o.Dval ** 2))
    comps = None
    if kind == "quadbatch":
        A_pat = sp.csr_matrix((np.ones_like(o.Aval), o.Acol, o.Arow), shape=(nf, nf))
        ncomp, comps = csg.connected_components(A_pat, directed=False)
    out = {"elements": int(info["nelem"]), "ndofs": int(nd), "nnz": int(len(o.Aval) + len(o.Dval))}
    for mode, tol in MODES.items():
        plan = prob.generate(mode)
        g = plan.csr(nd)
        for name in ("Arow", "Acol", "Drow", "Dcol"):
            assert np.array_equal(getattr(g, name), getattr(o, name)), (kind, mode, "pattern", name)
        ke = plan.element_matrices()[idx]
        r = {"pattern_bit_exact": True,
             "element_matrix": relerr(ke, ke_ref),
             "global": float(np.sqrt(np.sum((g.Aval - o.Aval) ** 2) + np.sum((g.Dval - o.Dval) ** 2)) / ref_glob),
             "rhs": relerr(g.rhs, o.rhs),
             "kernels": dict(plan.kernel_counts(mode), tile=plan.tile_count(mode))}
        if comps is not None:
            rows = np.repeat(np.arange(nf), np.diff(o.Arow))
            diff2 = np.bincount(comps[rows], weights=(g.Aval - o.Aval) ** 2, minlength=comps.max() + 1)
            ref2 = np.bincount(comps[rows], weights=o.Aval ** 2, minlength=comps.max() + 1)
            r["instances"] = int(len(ref2))
            r["per_instance_max"] = float(np.sqrt((diff2 / ref2).max()))
        print(kind, mode, r, flush=True)
        assert r["element_matrix"] <= tol and r["global"] <= tol and r["rhs"] <= tol, (kind, mode, r)
        if "per_instance_max" in r:
            assert r["per_instance_max"] <= tol
        out[mode] = r
    return out


def solution_errors(ctx, kind, spec):
    print("case parity solve", kind, spec, flush=True)
    prob = sa.Problem(ctx, kind, spec)
    info, o = oracle_system(prob)
    nd, nf = info["ndofs"], info["nfree"]
    A, D = csr_mats(o, nf, nd)
    cons = np.ctypeslib.as_array(prob.desc().constrained, shape=(nd,)).astype(bool)
    if not np.any(o.rhs):
        # no source term (the K + k^2 M systems): the solution is zero in every mode
        return {"ndofs": int(nd), "skipped": "zero right-hand side: the solve check is vacuous"}
    xd = o.rhs[cons]
    x_ref = np.zeros(nd)
    x_ref[cons] = xd
    x_ref[~cons] = spl.spsolve(A.tocsc(), o.rhs[~cons] - D @ xd)
    out = {"ndofs": int(nd)}
    for mode in MODES:
        plan = prob.generate(mode)
        x, it = plan.solve(nd, rel_tol=1e-12, max_outer=30, max_inner=20000)
        out[mode] = {"solution": relerr(x, x_ref), "rel_residual": it["rel_residual"], "method": it["method"]}
        print(kind, mode, out[mode], flush=True)
        assert it["rel_residual"] <= 1e-11
    assert out["fp64"]["solution"] <= 1e-8
    return out


@pytest.mark.parametrize("kind", list(CONFIGS))
def test_config_parity(ctx, kind):
    big, small = CONFIGS[kind]
    payload = {"spec": big, "matrix": matrix_errors(ctx, kind, big),
               "solve_spec": small, "solve": solution_errors(ctx, kind, small)}
    report(kind, payload)


def test_config_parity_newton(ctx):
    spec = "n=150;f=500"
    print("case parity newton2d", spec, flush=True)
    prob = sa.Problem(ctx, "newton2d", spec)
    info = prob.info()
    nd, nf = info["ndofs"], info["nfree"]
    cons = np.ctypeslib.as_array(prob.desc().constrained, shape=(nd,)).astype(bool)
    x = np.zeros(nd)
    steps = []
    for step in range(4):
        _, o = oracle_system(prob)
        ref_glob = np.sqrt(np.sum(o.Aval ** 2) + np.sum(o.Dval ** 2))
        row = {}
        for mode, tol in MODES.items():
            g = prob.generate(mode).csr(nd)
            assert np.array_equal(g.Acol, o.Acol) and np.array_equal(g.Arow, o.Arow)
            e = float(np.sqrt(np.sum((g.Aval - o.Aval) ** 2) + np.sum((g.Dval - o.Dval) ** 2)) / ref_glob)
            row[mode] = {"global": e, "rhs": relerr(g.rhs, o.rhs)}
            assert e <= tol and row[mode]["rhs"] <= tol, (step, mode, row[mode])
        A, _ = csr_mats(o, nf, nd)
        b = o.rhs[~cons]
        row["oracle_residual"] = float(np.linalg.norm(b - A @ x[~cons]) / np.linalg.norm(b))
        xn = x.copy()
        xn[~cons] = spl.spsolve(A.tocsc(), b)
        xn[cons] = o.rhs[cons]
        x = xn
        prob.setdata_vec(x)
        print("newton step", step, row, flush=True)
        steps.append(row)
    report("newton2d", {"spec": spec, "steps": steps})
