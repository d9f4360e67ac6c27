#!/usr/bin/env python3
"""Benchmark: element integration + global assembly on one MI355X.

Workload (default): plane-strain elasticity, 400 x 400 grid of T6 triangles
(320,000 elements, 1,283,202 dofs), 3-point rule, nodes jittered by
U(-0.15, 0.15) h (seeded SplitMix64), inputs resident in HBM.

A step is one assembly of the whole mesh into the global CSC matrix:
  * mixed  : REUSE method, FP32 element arithmetic, FP64 sums (the value);
  * fp64   : the plain FP64 method (gather, element kernel, COO in stream
             order, stable radix sort, in-order segmented reduce), every step;
  * fp64-reuse : REUSE method in FP64 (bit-identical to the plain method).
The pattern of the REUSE method is built once before the timed region and its
build time is reported separately.

Prints ONE JSON line.  --gpus N > 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "elfel.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=400, help="grid size of the T6 elasticity workload")
    ap.add_argument("--fp64-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-N", type=int, default=400)
    ap.add_argument("--only", choices=["all", "mixed", "fp64", "fp64-reuse"], default="all")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MiB",
                    help="only run a device copy of this size (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def rel_frob(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU per box; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import elfel_hip as eh
    from elfel_hip import _lib as L
    from elfel_hip import workloads as W

    t0 = time.time()
    cfg = W.elast_t6(args.N, frac=0.15)
    t_host = time.time() - t0
    fesp, n = cfg["fesp"], cfg["ndofs"]
    mesh = fesp.mesh
    E, Nn = mesh.nelements, mesh.nnodes
    ctx = L.Context(0)
    devname, ncu = ctx.device_name()
    t0 = time.time()
    ctx.set_mesh(0, mesh.conn, mesh.xy)
    ctx.set_dofs(0, fesp.dofnums, n)
    t_upload = time.time() - t0
    D = cfg["D"]
    rule = 3

    def run(precision, method, steps, warmup):
        ctx.timing(False)
        nnz = ctx.assemble(L.ELAST, L.T6, rule, precision=precision, method=method, layout=L.CSC, D=D)
        ctx.synchronize()
        for _ in range(warmup):
            ctx.assemble(L.ELAST, L.T6, rule, precision=precision, method=method, layout=L.CSC, D=D, wait=False)
        ctx.synchronize()
        ctx.timing(True)
        w0 = time.perf_counter()
        for _ in range(steps):
            ctx.assemble(L.ELAST, L.T6, rule, precision=precision, method=method, layout=L.CSC, D=D, wait=False)
        ctx.synchronize()
        wall = (time.perf_counter() - w0) * 1e3 / steps
        ph = ctx.timing_read()
        ctx.timing(False)
        ptr, ind, val, shape, lay = ctx.copy_sparse(base=0)
        return dict(nnz=nnz, wall_ms=wall, phases=ph, val=val, ptr=ptr, ind=ind)

    res = {}
    # Symbolic phase of the REUSE method (pattern + tiles), timed once.
    ctx.timing(True)
    t0 = time.perf_counter()
    ctx.assemble(L.ELAST, L.T6, rule, precision=L.MIXED, method=L.REUSE, layout=L.CSC, D=D)
    ctx.synchronize()
    t_sym = (time.perf_counter() - t0) * 1e3
    ctx.timing(False)
    if args.only in ("all", "mixed"):
        res["mixed"] = run(L.MIXED, L.REUSE, args.steps, args.warmup)
    if args.only in ("all", "fp64-reuse"):
        res["fp64_reuse"] = run(L.FP64, L.REUSE, args.steps, args.warmup)
    if args.only in ("all", "fp64"):
        res["fp64"] = run(L.FP64, L.SORT, max(1, min(args.fp64_steps, args.steps)), min(2, args.warmup))
    mindet, nfallback = ctx.geometry_stats()

    nnz = next(iter(res.values()))["nnz"]
    # algorithmic bytes
    survey_bytes = E * 6 * 4 + Nn * 16 + n * 4 + nnz * (8 + 4) + (n + 1) * 4  # full assembly, mesh -> CSC
    geom_bytes = E * 6 * 2 * 8                              # FP64 element coordinate records
    kernel_bytes_mixed = geom_bytes + nnz * 8               # k_reuse: element records in, values out
    step_bytes_reuse = E * 6 * 4 + Nn * 16 + nnz * 8       # gather + k_reuse: conn + coords in, values out

    out = {}
    for key, r in res.items():
        ph = r["phases"]
        steps = max(1, ph["total"][1])
        ms = ph["total"][0] / steps
        out[key] = dict(ms_per_step=ms, wall_ms_per_step=r["wall_ms"], elements_per_s=E / (ms * 1e-3),
                        numeric_ms=(ph["numeric"][0] / ph["numeric"][1]) if ph["numeric"][1] else None,
                        gather_ms=(ph["gather"][0] / ph["gather"][1]) if ph["gather"][1] else None,
                        element_ms=(ph["element"][0] / ph["element"][1]) if ph["element"][1] else None,
                        sort_ms=(ph["sort"][0] / ph["sort"][1]) if ph["sort"][1] else None,
                        reduce_ms=(ph["reduce"][0] / ph["reduce"][1]) if ph["reduce"][1] else None,
                        survey_gbs=survey_bytes / (ms * 1e-3) / 1e9)
    errs = {}
    if "fp64" in res and "mixed" in res:
        assert np.array_equal(res["fp64"]["ptr"], res["mixed"]["ptr"]) and np.array_equal(res["fp64"]["ind"],
                                                                                           res["mixed"]["ind"])
        errs["mixed_vs_fp64_frobenius"] = rel_frob(res["mixed"]["val"], res["fp64"]["val"])
        errs["mixed_vs_fp64_max_entry_over_max"] = float(np.abs(res["mixed"]["val"] - res["fp64"]["val"]).max() /
                                                         np.abs(res["fp64"]["val"]).max())
    if "fp64" in res and "fp64_reuse" in res:
        errs["fp64_reuse_equals_fp64_plain"] = bool(np.array_equal(res["fp64"]["val"], res["fp64_reuse"]["val"]))

    m = out.get("mixed") or next(iter(out.values()))
    kern_ms = m["numeric_ms"] or m["ms_per_step"]
    achieved = kernel_bytes_mixed / (kern_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_reuse<ELAST, float, T6, CSC>", "kernel_ms": kern_ms,
                "algorithmic_bytes": kernel_bytes_mixed}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_sample_N)

    line = {
        "metric": "elements assembled/s + HBM GB/s: FP32-mixed vs FP64 GPU path at <=1e-6 rel err",
        "value": m["elements_per_s"],
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 element arithmetic, f64 sums",
        "data": "synthetic (seeded SplitMix64 jitter), inputs resident in HBM",
        "config": {"workload": f"plane-strain elasticity T6 {args.N}x{args.N}, 3-pt rule, +-15% h node jitter",
                   "elements": E, "nodes": Nn, "dofs": n, "nnz": nnz, "layout": "CSC (finish! layout)",
                   "method": "REUSE (pattern cached, fused owner-computes kernel)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "paths": out,
        "speedup_mixed_vs_fp64_plain": (out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"])
        if ("fp64" in out and "mixed" in out) else None,
        "step_gbs_mixed": step_bytes_reuse / (m["ms_per_step"] * 1e-3) / 1e9,
        "errors": errs,
        "geometry": {"min_det": mindet, "fp64_branch_elements": nfallback},
        "setup": {"host_mesh_and_numbering_s": t_host, "upload_s": t_upload, "symbolic_ms": t_sym},
        "device": {"name": devname, "cus": ncu},
    }
    print(json.dumps(line))


def copy_probe(mib):
    """A device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    counters give the scale factor between counter units and bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": n * 4, "launches": 3}))


def cpu_baseline(N):
    """The oracle (C, one thread) on the same workload: element loop + sparse()."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
    except Exception as e:  # the checker is optional for the GPU numbers
        return {"error": f"oracle unavailable: {e}"}
    from elfel_hip import workloads as W
    cfg = W.elast_t6(N, frac=0.15)
    fesp = cfg["fesp"]
    p = O.make_problem(O.ELAST, O.T6, 3, fesp.mesh.conn, fesp.mesh.xy, fesp.dofnums, D=cfg["D"])
    te, ts, nnz = O.time_assembly(p, reps=1, nthreads=1)
    E = fesp.mesh.nelements
    return {"value": E / (te + ts), "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": f"full {N}x{N} T6 mesh ({E} elements), 1 repetition: element loop {te:.3f} s + "
                      f"sparse() {ts:.3f} s, single thread (the reference's path is serial)",
            "element_loop_s": te, "sparse_s": ts, "host_cores_visible": os.cpu_count()}


if __name__ == "__main__":
    main()
