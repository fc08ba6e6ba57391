#!/usr/bin/env python3
"""Benchmark: assembled nnz/s of the mixed-precision MFMA path vs the plain FP64
GPU path on the 3D linear-elasticity workload (h1xyz P2 on 64^3 x 6 Kuhn
tetrahedra, jittered, E lognormal per element, nu = 0.3).

One step = one numeric assembly (Jacobians, Gauss-point coefficients, element
contractions, deterministic FP64 reduction into the cached CSR pattern + rhs)
with the node coordinates and parameters already resident in HBM.

Prints one JSON line (see the keys below).  One GPU only: --gpus > 1 exits.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {  # MI355X published dense peaks
    "hbm_GBps": 8000.0,
    "fp16_TFLOPs": 2500.0,
    "bf16_TFLOPs": 2500.0,
    "fp32_TFLOPs": 157.3,
    "fp64_TFLOPs": 78.6,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def element_flops(c, n, d, ngp):
    """Canonical B^T D B contraction flops per element: 2 n_gp (c n)^2 (c d)."""
    return 2.0 * ngp * (c * n) ** 2 * (c * d)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prewarm-ms", type=float, default=250.0,
                    help="untimed steps of each mode filling this much wall time before its W warmup steps "
                         "(a sub-millisecond step leaves the GPU at its idle clocks otherwise)")
    ap.add_argument("--problem", default="elasticity3d",
                    help="elasticity3d (the headline workload), hppoisson3d (hp hexahedra, orders 1..5), poisson2d, "
                         "quadbatch (512 independent problems), newton2d (Newton re-assembly with a field), prism3d "
                         "(order-1 prisms) or hcurlprojection (hcurl fields; --extra ';elem=tet;order=2;term=maxwell')")
    ap.add_argument("--n", type=int, default=0,
                    help="cells per edge (0: the workload's full size: 64 elasticity3d, 48 hppoisson3d, 200 poisson2d, "
                         "708 newton2d, 96 prism3d; quadbatch takes its sizes from --extra)")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--extra", default="", help="extra problem options, e.g. ';pmin=5;pmax=5'")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run (timing sweeps)")
    ap.add_argument("--mixed", default="f16split", help="mixed precision mode of the headline value")
    ap.add_argument("--variants", default="bf16split,fp32mfma,fp32",
                    help="further precision modes timed in the same run and reported under 'variants' ('' for none)")
    ap.add_argument("--cpu-sample-n", type=int, default=0, help="CPU baseline sample size (0: auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the device solve of the assembled system (slx_plan_solve) reported under 'solve'")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % args.gpus)
        sys.exit(2)

    import numpy as np
    import torch
    import sparselizard_amd as sa

    dev = torch.device("cuda:0")
    torch.cuda.init()
    if args.copy_probe_mib > 0:
        copy_probe(torch, dev, args.copy_probe_mib)
        return
    ctx = sa.Context(0)
    if args.n <= 0:
        args.n = {"elasticity3d": 64, "hppoisson3d": 48, "poisson2d": 200, "newton2d": 708, "prism3d": 96,
                  "hcurlprojection": 32}.get(args.problem, 0)
    if args.problem == "elasticity3d":
        spec = "n=%d;order=%d" % (args.n, args.order)
    elif args.problem == "quadbatch":
        spec = ""
    else:
        spec = "n=%d" % args.n
    spec += args.extra
    t0 = time.time()
    prob = sa.Problem(ctx, args.problem, spec)
    info = prob.info()
    t_host = time.time() - t0
    log("problem built in %.1f s: %s" % (t_host, info))
    t0 = time.time()
    plan = prob.plan()
    t_sym = time.time() - t0
    shape = plan.shape()
    stats = plan.stats()
    nnz = shape["nnzA"] + shape["nnzD"]
    log("plan (symbolic) in %.1f s: %s %s" % (t_sym, shape, stats))

    nd = prob.numeric()
    nn = info["ndofs"]
    # device-resident inputs
    import ctypes as C
    xyz_host = np.ctypeslib.as_array(nd.node_xyz, shape=(3 * prob.n_nodes(),)).copy()
    par_host = (np.ctypeslib.as_array(nd.params, shape=(nd.n_params * nd.param_len,)).copy()
                if nd.n_params > 0 and nd.param_len > 0 else np.zeros(1))
    d_xyz = torch.from_numpy(xyz_host).to(dev)
    d_par = torch.from_numpy(par_host).to(dev)
    # field data read by the coefficient programs (Newton workloads), device-resident
    d_fdata = [torch.from_numpy(f.copy()).to(dev) for f in prob.field_data()]
    d_fields = (C.c_void_p * len(d_fdata))(*[f.data_ptr() for f in d_fdata]) if d_fdata else None
    stream = torch.cuda.current_stream()
    sptr = C.c_void_p(stream.cuda_stream)

    def step(prec):
        plan.assemble_device(C.c_void_p(d_xyz.data_ptr()), C.c_void_p(d_par.data_ptr()), nd.param_len, d_fields, prec, sptr)

    def timed(prec):
        t_pre, n_pre = time.time(), 0
        while (time.time() - t_pre) * 1e3 < args.prewarm_ms and n_pre < 20000:
            step(prec)
            n_pre += 1
            if n_pre % 8 == 0:
                torch.cuda.synchronize()
        for _ in range(args.warmup):
            step(prec)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t_wall = time.time()
        ev0.record(stream)
        for _ in range(args.steps):
            step(prec)
        ev1.record(stream)
        torch.cuda.synchronize()
        t_wall = time.time() - t_wall
        return ev0.elapsed_time(ev1) / args.steps, t_wall

    def csr_now():
        return plan.csr(nn)

    # plain FP64 GPU path
    if args.no_fp64:
        ms64, ref = float("nan"), None
    else:
        ms64, _ = timed("fp64")
        ref = csr_now()
        log("fp64: %.3f ms/step" % ms64)
    # mixed precision path (the headline)
    msmix, wall = timed(args.mixed)
    mix = csr_now()
    log("%s: %.3f ms/step" % (args.mixed, msmix))

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    def glob_err(m):
        return float(np.sqrt(np.sum((m.Aval - ref.Aval) ** 2) + np.sum((m.Dval - ref.Dval) ** 2)) /
                     np.sqrt(np.sum(ref.Aval ** 2) + np.sum(ref.Dval ** 2)))

    if ref is None:
        err_A = err_D = err_b = glob = float("nan")
    else:
        err_A = rel(mix.Aval, ref.Aval)
        err_D = rel(mix.Dval, ref.Dval) if ref.Dval.size else 0.0
        err_b = rel(mix.rhs, ref.rhs)
        glob = glob_err(mix)
    del mix

    # the other precision modes on the same plan and inputs
    variants = {}
    for mode in [v for v in args.variants.split(",") if v and v != args.mixed]:
        ms_v, _ = timed(mode)
        entry = {"ms_per_step": ms_v, "nnz_per_s": nnz / (ms_v * 1e-3), "ratio_vs_fp64": ms64 / ms_v,
                 "kernels": plan.kernel_counts(mode)}
        if ref is not None:
            vcsr = csr_now()
            entry["rel_err_vs_fp64_gpu"] = {"global_frobenius": glob_err(vcsr), "rhs": rel(vcsr.rhs, ref.rhs)}
            del vcsr
        entry["kernel_ms"] = plan.profile(C.c_void_p(d_xyz.data_ptr()), C.c_void_p(d_par.data_ptr()), nd.param_len,
                                          d_fields, mode, iters=3)
        log("%s: %.3f ms/step %s" % (mode, ms_v, entry.get("rel_err_vs_fp64_gpu", "")))
        variants[mode] = entry
    if variants:
        # leave the headline mode's matrix in the plan (the solve below uses it)
        step(args.mixed)
        torch.cuda.synchronize()

    # per-kernel timing: HIP events around every launch on the plan's stream
    prof_mix = plan.profile(C.c_void_p(d_xyz.data_ptr()), C.c_void_p(d_par.data_ptr()), nd.param_len, d_fields,
                            args.mixed, iters=max(3, args.steps // 2))
    prof_64 = None if args.no_fp64 else plan.profile(C.c_void_p(d_xyz.data_ptr()), C.c_void_p(d_par.data_ptr()), nd.param_len,
                                                      d_fields, "fp64", iters=max(3, args.steps // 2))
    if args.problem == "elasticity3d":
        n_elem = count_tets(args.n)
        n = (args.order + 1) * (args.order + 2) * (args.order + 3) // 6
        ngp = {1: 11, 2: 24, 3: 43}.get(args.order, 24)
        flops = element_flops(3, n, 3, ngp) * n_elem
    else:
        # elements of the matrix buckets (every element once; rhs buckets repeat them)
        dd = prob.desc()
        n_elem = sum(dd.buckets[b].n_elem for b in range(dd.n_buckets)
                     if any(not dd.buckets[b].slices[s].is_rhs for s in range(dd.buckets[b].n_slices)))
        flops = slice_flops(dd)
    ke_bytes = 8 if args.mixed == "fp64" else 4
    # reduction, compulsory bytes (the per-unit figure of the bytes formula):
    # every CSR value written once; the element-matrix round trip through HBM
    # (written by the contraction, read here) is not compulsory and is
    # reported apart, as is the slot metadata
    red_bytes = nnz * 8
    ke_roundtrip = stats["n_coo"] * ke_bytes
    mixed_kind = {"f16split": "fp16", "bf16split": "bf16", "fp32mfma": "fp32", "fp32": "fp32"}.get(args.mixed, "fp32")
    peak_elem = PEAKS[mixed_kind + "_TFLOPs"]
    t_contract = prof_mix["mfma"] if prof_mix["mfma"] > 0 else prof_mix["valu"]
    # contraction kernel: canonical flops, and its compulsory HBM bytes (the
    # coefficient table read, every element-matrix entry written once)
    d = prob.desc()
    con_bytes = 0.0
    for b in range(d.n_buckets):
        bd = d.buckets[b]
        nq = 1 if bd.q_invariant else bd.n_gp
        con_bytes += bd.n_coefs * nq * bd.n_elem * 4.0
        for s_ in range(bd.n_slices):
            sd = bd.slices[s_]
            if not sd.is_rhs:
                con_bytes += bd.n_elem * sd.n_tf * sd.n_dof * 4.0
    kname = ("factor_contract_kernel" if args.problem == "hppoisson3d" else "mfma_contract_kernel") \
        if prof_mix["mfma"] > 0 else "elem_cols"
    if t_contract >= prof_mix["reduce"]:
        # the roofline that bounds the kernel: the larger of the two ideal times
        t_flop = flops / (peak_elem * 1e12)
        t_byte = con_bytes / (PEAKS["hbm_GBps"] * 1e9)
        if t_flop >= t_byte:
            achieved = flops / (t_contract * 1e-3) / 1e12
            roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak_elem, "unit": "TFLOP/s",
                    "frac": round(achieved / peak_elem, 5), "traffic": None, "kernel": kname,
                    "kernel_ms": round(t_contract, 4), "algorithmic": flops}
        else:
            achieved = con_bytes / (t_contract * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAKS["hbm_GBps"], "unit": "GB/s",
                    "frac": round(achieved / PEAKS["hbm_GBps"], 5), "traffic": None, "kernel": kname,
                    "kernel_ms": round(t_contract, 4), "algorithmic": con_bytes,
                    "mfma_tflops": round(flops / (t_contract * 1e-3) / 1e12, 3)}
    else:
        achieved = red_bytes / (prof_mix["reduce"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAKS["hbm_GBps"], "unit": "GB/s",
                "frac": round(achieved / PEAKS["hbm_GBps"], 5), "traffic": None, "kernel": "reduce_rows",
                "kernel_ms": round(prof_mix["reduce"], 4), "algorithmic": red_bytes,
                "ke_roundtrip_bytes": ke_roundtrip}
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        try:
            tr = json.load(open(traffic_file))
            key = "%s|%s|n%d" % (roof["kernel"], args.mixed, args.n)
            if key in tr:
                roof["traffic"] = tr[key]
        except (OSError, ValueError):
            pass

    solve = None
    if not args.no_solve:
        # sl::solve on the last assembled system: FP64 refinement around FP32 Jacobi-PCG
        torch.cuda.synchronize()
        t0 = time.time()
        x, it = plan.solve(nn, rel_tol=1e-10, max_outer=20, max_inner=20000)
        t_solve = time.time() - t0
        # per inner iteration: FP32 values + column indices once, p gathered
        # (at least once), q, d, r, z, p, diag streamed
        it_bytes = shape["nnzA"] * 8 + shape["nfree"] * 4 * 11
        solve = {"s": t_solve, "method": it["method"], "outer": it["outer"], "inner": it["inner"], "rel_residual": it["rel_residual"],
                 "ms_per_inner": t_solve * 1e3 / max(it["inner"], 1), "bytes_per_inner": it_bytes,
                 "GBps_per_inner": it_bytes / (t_solve / max(it["inner"], 1)) / 1e9, "x_norm": float(np.linalg.norm(x))}
        log("solve: %s" % solve)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sa, args)

    out = {
        "metric": "assembled nnz/s: MFMA-batched mixed-prec vs plain FP64 GPU at rel err <=1e-6",
        "value": nnz / (msmix * 1e-3),
        "unit": "nnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "prewarm_ms": args.prewarm_ms,
        "ms_per_step": msmix,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f16split": "f16x2-split/f32acc/f64reduce", "bf16split": "bf16x3-split/f32acc/f64reduce",
                  "fp32": "f32/f64reduce"}.get(args.mixed, args.mixed),
        "data": "synthetic",
        "config": {"workload": workload_name(args, n_elem),
                   "n_elem": n_elem, "ndofs": nn, "nnzA": shape["nnzA"], "nnzD": shape["nnzD"],
                   "coo_entries": stats["n_coo"], "parallelism": "single GPU"},
        "fp64_gpu": {"ms_per_step": ms64, "nnz_per_s": nnz / (ms64 * 1e-3), "kernel_ms": prof_64},
        "mixed_gpu": {"mode": args.mixed, "ms_per_step": msmix, "kernel_ms": prof_mix},
        "ratio_mixed_vs_fp64": ms64 / msmix,
        "rel_err_vs_fp64_gpu": {"global_frobenius": glob, "A": err_A, "D": err_D, "rhs": err_b},
        "symbolic_s": t_sym,
        "roofline": roof,
        "cpu_baseline": cpu,
        "variants": variants,
    }
    if solve is not None:
        out["solve"] = solve
    print(json.dumps(out))


def workload_name(args, n_elem):
    if args.problem == "elasticity3d":
        return ("3D linear elasticity h1xyz P%d, %d^3 x 6 jittered tets (%d), E lognormal(0,0.5) per element, nu=0.3, "
                "clamped x=0" % (args.order, args.n, n_elem))
    if args.problem == "hppoisson3d":
        return ("3D hp-Poisson, %d^3 jittered hexahedra (%d), per-element order U{1..5}, c lognormal(0,1), Gauss order 2p+1"
                % (args.n, n_elem))
    if args.problem == "poisson2d":
        return "2D Poisson P1, %dx%dx2 jittered triangles (%d), f=1, u=0 on the boundary" % (args.n, args.n, n_elem)
    if args.problem == "quadbatch":
        return ("512 independent K + k^2 M problems, n x n quads of order 3, n in [16,160], k^2 in [1,50] (%d elements)%s"
                % (n_elem, args.extra))
    if args.problem == "newton2d":
        return ("2D nonlinear magnetostatics Newton re-assembly, P2, %dx%dx2 jittered triangles (%d), "
                "nu(|grad a|) = 1 + 500/(1+|grad a|^2)" % (args.n, args.n, n_elem))
    if args.problem == "prism3d":
        order = 1
        for kv in args.extra.split(";"):
            if kv.startswith("order="):
                order = int(kv[6:])
        return ("3D Poisson + mass, order-%d prisms, %d^3 x 2 jittered prisms (%d), c lognormal(0,1), u=0 on the side faces"
                % (order, args.n, 2 * args.n ** 3))
    if args.problem == "hcurlprojection":
        opts = dict(kv.split("=", 1) for kv in args.extra.split(";") if "=" in kv)
        elem = opts.get("elem", "tri")
        cells = "%d^3" % args.n if elem in ("tet", "hex") else "%dx%d" % (args.n, args.n)
        return ("hcurl %s of order %s on %s %s cells (%d elements, nodes randomly renumbered): %s"
                % (opts.get("term", "mass"), opts.get("order", "1"), cells, elem, n_elem,
                   "integral(-curl(dof(E)) curl(tf(E)) + k2 dof(E) tf(E) - f tf(E))" if opts.get("term") == "maxwell"
                   else "hcurl formulation"))
    return "%s %s" % (args.problem, args.extra)


def copy_probe(torch, dev, mib):
    """A plain device-to-device copy of a known size: the FETCH_SIZE and
    WRITE_SIZE counters of its kernel give the scale of those counters."""
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4}))


def slice_flops(d):
    """Canonical contraction flops of every matrix slice, sum over buckets:
    2 n_gp n_tf n_dof n_a per element (n_a: tf derivatives of the slice's
    terms), which is 6 (p+1)^9 for a scalar Poisson hexahedron of order p."""
    total = 0.0
    for b in range(d.n_buckets):
        bd = d.buckets[b]
        for s in range(bd.n_slices):
            sd = bd.slices[s]
            if sd.is_rhs:
                continue
            na = len({sd.terms[t].tf_deriv for t in range(sd.n_terms)})
            total += 2.0 * bd.n_gp * sd.n_tf * sd.n_dof * na * bd.n_elem
    return total


def count_tets(n):
    return 6 * n ** 3


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sa, args):
    """The CPU oracle (restated generate()+A()+b(), FP64) on bounded samples of
    the same workload, timed on this host: one thread (the reference's generate
    loop is single-threaded), and all the host cores this process may use (16
    at most: the share of one GPU), where the oracle threads the element loop
    and, like rawmat::process (rawmat.cpp:221), the per-row sort.  nnz/s is
    reported for each sample; 'value' is the all-cores figure."""
    from oracle import py_oracle

    # sample size knob, its start, cap and how the element count grows with it
    knob = {"elasticity3d": ("n", 4, 28, 3), "hppoisson3d": ("n", 3, 24, 3), "poisson2d": ("n", 20, 1200, 2),
            "newton2d": ("n", 20, 708, 2), "quadbatch": ("instances", 2, 96, 1),
            "prism3d": ("n", 4, 110, 3),
            "hcurlprojection": ("n", 3, 40, 3) if any(e in args.extra for e in ("elem=tet", "elem=hex")) else ("n", 4, 400, 2)
            }[args.problem]

    def spec_of(v):
        if args.problem == "elasticity3d":
            return "n=%d;order=%d" % (v, args.order)
        return "%s=%d%s" % (knob[0], v, args.extra)

    def run(v, nthreads):
        p = sa.Problem(None, args.problem, spec_of(v))
        info = p.info()
        nd = p.numeric()
        t0 = time.time()
        o = py_oracle.assemble(p.desc_ptr(), nd, info["ndofs"], info["nfree"], nthreads=nthreads)
        dt = time.time() - t0
        return dt, len(o.Aval) + len(o.Dval), info["nelem"]

    def sized(nthreads, seconds):
        v = args.cpu_sample_n
        if v <= 0:
            dt, nnz, ne = run(knob[1], nthreads)
            grow = seconds / max(dt, 1e-9)
            v = max(knob[1], int(round(knob[1] * grow ** (1.0 / knob[3]))))
            v = min(v, knob[2])
        dt, nnz, ne = run(v, nthreads)
        return {"value": nnz / dt, "s": dt, "sample": "%s with %s (%d elements)" % (args.problem, spec_of(v), ne)}

    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = max(1, min(16, avail))
    one = sized(1, 10.0)
    allc = sized(cores, 10.0)
    return {"value": allc["value"], "unit": "nnz/s", "cores": cores, "kind": "port",
            "sample": "oracle generate()+A()+b() FP64, %d threads, %s, %.1f s" % (cores, allc["sample"], allc["s"]),
            "single_thread": {"value": one["value"], "cores": 1, "sample": "%s, %.1f s" % (one["sample"], one["s"])},
            "cpu_model": cpu_model()}


if __name__ == "__main__":
    main()
