#!/usr/bin/env python3
"""Benchmark: element integration + global assembly on one MI355X.

Default workload (--config c5), the largest configuration and the one that
carries the solve: heat on a 4096 x 4096 grid of Q4 quadrilaterals
(16,777,216 elements, 16,785,409 dofs), 2x2 Gauss, conductivity
lognormal(0, 1) and source U[-1, 1] per element (seeded SplitMix64), u = 0 on
the boundary, inputs resident in HBM, CSR output + load vector; then
Jacobi-PCG on the 16,769,025 free dofs to ||b - A x|| / ||b|| <= 1e-10, FP32
iterations with FP64 residual replacement against FP64 PCG (the "pcg" object).

A step is one assembly of the whole mesh into the global sparse matrix:
  * mixed      : REUSE method, FP64 geometry, FP32 element products, FP64 sums
                 (the value);
  * fp64       : the plain FP64 method (element kernel, COO in stream order,
                 stable radix sort, in-order segmented reduce), every step;
  * fp64-reuse : REUSE method in FP64 (bit-identical to the plain method).
The pattern of the REUSE method is built once before the timed region and its
build time is reported separately.

Other workloads, each printing its own line:
  c1  heat T3 100x100, 1-point rule, CSR + load vector;
  c2  plane-strain elasticity T6 400x400 (320,000 elements, 1,283,202 dofs),
      3-point rule, nodes jittered by U(-0.15, 0.15) h, CSC;
  c3  Stokes Taylor-Hood T6/T3 500x500, K and B blocks in one launch, CSC;
  c4  256 heat-T3 / elasticity-T6 problems of uneven size in one persistent
      launch replayed from a hipGraph (mixed), against the FP64 plain method
      problem by problem.
Element families (the REUSE method with a cached sort permutation, mixed:
FP32 element values, FP64 sums; beside it the FP64 REUSE method, bit for bit
the plain method, and the plain FP64 method):
  t4  heat on a 64^3 block of T4 tetrahedra (examples/heat/poisson/t4.jl);
  p1b Reddy-form Stokes, T3 + bubble velocity / T3 pressure, 512 x 512
      (examples/stokes/colliding_flow/p1b_p1.jl).

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
METRIC = "elements assembled/s + HBM GB/s: FP32-mixed vs FP64 GPU path at <=1e-6 rel err"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", choices=["c1", "c2", "c3", "c4", "c5", "t4", "p1b"], default="c5")
    ap.add_argument("--N", type=int, default=0, help="grid size (default: the configuration's)")
    ap.add_argument("--fp64-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pcg", action="store_true", help="c5: assembly only")
    ap.add_argument("--pcg-maxit", type=int, default=60000)
    ap.add_argument("--only", choices=["all", "mixed", "fp64", "fp64-reuse"], default="all")
    ap.add_argument("--kinds", choices=["all", "heat", "elast"], default="all",
                    help="c4: only the problems of this kind (a diagnostic split of the batch)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MiB",
                    help="only run a device copy of this size (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def rel_frob(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------------------
# single-problem workloads
# ---------------------------------------------------------------------------
def setup_case(cfgname, N, ctx, L, W, cfg=None):
    """Uploads the mesh(es) and dof maps; returns the assemble() arguments and
    the sizes that enter the byte counts."""
    if cfgname == "c1":
        N = N or 100
        cfg = cfg or W.heat_t3(N)
        f = cfg["fesp"]
        ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
        ctx.set_dofs(0, f.dofnums, cfg["ndofs"])
        return dict(cfg=cfg, N=N, kind=L.HEAT, fe=L.T3, rule=1, layout=L.CSR, kw=dict(kappa=1.0, Q=-6.0),
                    workload=f"heat T3 {N}x{N}, 1-pt rule, kappa 1, Q -6, CSR + load vector",
                    E=f.mesh.nelements, Nn=f.mesh.nnodes, n=cfg["ndofs"], nen=3, nen_p=0, ncoef=0, heat=True,
                    kernel="k_reuse<HEAT, float, T3, CSR>", oracle=dict(kind="heat", fesp=f))
    if cfgname == "c2":
        N = N or 400
        cfg = cfg or W.elast_t6(N, frac=0.15)
        f = cfg["fesp"]
        ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
        ctx.set_dofs(0, f.dofnums, cfg["ndofs"])
        return dict(cfg=cfg, N=N, kind=L.ELAST, fe=L.T6, rule=3, layout=L.CSC, kw=dict(D=cfg["D"]),
                    workload=f"plane-strain elasticity T6 {N}x{N}, 3-pt rule, +-15% h node jitter",
                    E=f.mesh.nelements, Nn=f.mesh.nnodes, n=cfg["ndofs"], nen=6, nen_p=0, ncoef=0, heat=False,
                    kernel="k_reuse<ELAST, float, T6, CSC>", oracle=dict(kind="elast", fesp=f, D=cfg["D"]))
    if cfgname == "c3":
        N = N or 500
        cfg = cfg or W.stokes_t6t3(N)
        u, p = cfg["ufesp"], cfg["pfesp"]
        n = cfg["ndofs"]
        ctx.set_mesh(0, u.mesh.conn, u.mesh.xy)
        ctx.set_dofs(0, u.dofnums, n)
        ctx.set_mesh(1, p.mesh.conn, p.mesh.xy)
        ctx.set_dofs(1, p.dofnums, n)
        return dict(cfg=cfg, N=N, kind=L.STOKES_GEN, fe=L.T6, rule=3, layout=L.CSC, kw=dict(D=cfg["D"], mu=1.0),
                    workload=f"Stokes Taylor-Hood T6/T3 {N}x{N}, 3-pt rule, K and B in one launch",
                    E=u.mesh.nelements, Nn=u.mesh.nnodes, n=n, nen=6, nen_p=3, ncoef=0, heat=False,
                    kernel="k_reuse<STOKES_GEN, float, T6, CSC>",
                    oracle=dict(kind="stokes", fesp=u, pfesp=p, D=cfg["D"]))
    if cfgname == "c5":
        N = N or 4096
        cfg = cfg or W.heat_q4(N, sigma=1.0)
        f = cfg["fesp"]
        ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
        ctx.set_dofs(0, f.dofnums, cfg["ndofs"])
        ctx.set_coefficients(f.mesh.nelements, cfg["kappa_e"], cfg["Q_e"], None)
        return dict(cfg=cfg, N=N, kind=L.HEAT, fe=L.Q4, rule=2, layout=L.CSR, kw=dict(),
                    workload=f"heat Q4 {N}x{N}, 2x2 Gauss, kappa_e ~ lognormal(0,1), f_e ~ U[-1,1], CSR + load vector",
                    E=f.mesh.nelements, Nn=f.mesh.nnodes, n=cfg["ndofs"], nen=4, nen_p=0, ncoef=2, heat=True,
                    kernel="k_reuse<HEAT, float, Q4, CSR>",
                    oracle=dict(kind="heat", fesp=f, kappa_e=cfg["kappa_e"], Q_e=cfg["Q_e"]))
    raise ValueError(cfgname)


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_single(args, L, W):
    ctx = L.Context(0)
    devname, ncu = ctx.device_name()
    t0 = time.time()
    cs = setup_case(args.config, args.N, ctx, L, W)
    t_setup = time.time() - t0
    progress(f"{args.config}: mesh and numbering ready ({t_setup:.1f} s)")
    E, Nn, n = cs["E"], cs["Nn"], cs["n"]

    def assemble(precision, method, wait=True):
        return ctx.assemble(cs["kind"], cs["fe"], cs["rule"], precision=precision, method=method,
                            layout=cs["layout"], wait=wait, **cs["kw"])

    def run(precision, method, steps, warmup):
        ctx.timing(False)
        nnz = assemble(precision, method)
        ctx.synchronize()
        ramp = clock_ramp(lambda: assemble(precision, method, wait=False), ctx.synchronize)
        for _ in range(warmup):
            assemble(precision, method, wait=False)
        # the timed steps: device idle and synchronised on both sides, no events inside
        ctx.synchronize()
        w0 = time.perf_counter()
        for _ in range(steps):
            assemble(precision, method, wait=False)
        ctx.synchronize()
        wall = (time.perf_counter() - w0) * 1e3 / steps
        # the same steps again between one pair of hipEvents on the library's
        # stream: the average step (one fused kernel for the reuse method)
        ctx.timing_region()
        for _ in range(steps):
            assemble(precision, method, wait=False)
        rg = ctx.timing_read()["total"]
        region = rg[0] / max(1, rg[1])
        # and again with hipEvents around every phase (per-kernel times)
        ctx.timing(True)
        for _ in range(steps):
            assemble(precision, method, wait=False)
        ctx.synchronize()
        ph = ctx.timing_read()
        ctx.timing(False)
        ptr, ind, val, shape, lay = ctx.copy_sparse(base=0)
        F = ctx.copy_vector(n) if cs["heat"] else None
        return dict(nnz=nnz, wall_ms=wall, region_ms=region, phases=ph, val=val, ptr=ptr, ind=ind, F=F, ramp=ramp)

    res = {}
    # The library's code objects load at the first launch of each of their
    # kernels, once per process: a small mesh of the same problem in a context
    # of its own loads them, so that the first call below is that of a new
    # mesh, not of a new process.
    wctx = L.Context(0)
    wcs = setup_case(args.config, 8, wctx, L, W)
    for prec, meth in ((L.MIXED, L.REUSE), (L.FP64, L.REUSE), (L.FP64, L.SORT)):
        wctx.assemble(wcs["kind"], wcs["fe"], wcs["rule"], precision=prec, method=meth, layout=wcs["layout"], **wcs["kw"])
    wctx.synchronize()
    wctx.close()
    # The FP64 plain method's first call in a context of its own (its
    # allocations included), for the first-call comparison
    pctx = L.Context(0)
    pcs = setup_case(args.config, args.N, pctx, L, W, cfg=cs["cfg"])
    pctx.synchronize()
    t0 = time.perf_counter()
    pctx.assemble(pcs["kind"], pcs["fe"], pcs["rule"], precision=L.FP64, method=L.SORT, layout=pcs["layout"],
                  **pcs["kw"])
    pctx.synchronize()
    t_plain_first = (time.perf_counter() - t0) * 1e3
    pctx.close()
    del pcs
    # Symbolic phase of the REUSE method (pattern + tiles) and the numeric
    # kernel: the first mixed call on this context and mesh, timed once.
    t0 = time.perf_counter()
    assemble(L.MIXED, L.REUSE)
    ctx.synchronize()
    t_sym = (time.perf_counter() - t0) * 1e3
    progress(f"first REUSE call {t_sym:.1f} ms (FP64 plain method's first call {t_plain_first:.1f} ms)")
    if args.only in ("all", "mixed"):
        res["mixed"] = run(L.MIXED, L.REUSE, args.steps, args.warmup)
        progress("mixed reuse timed")
    if args.only in ("all", "fp64-reuse"):
        res["fp64_reuse"] = run(L.FP64, L.REUSE, args.steps, args.warmup)
        progress("fp64 reuse timed")
    if args.only in ("all", "fp64"):
        res["fp64"] = run(L.FP64, L.SORT, max(1, min(args.fp64_steps, args.steps)), min(2, args.warmup))
        progress("fp64 plain timed")
    mindet, nbad = ctx.geometry_stats()

    nnz = next(iter(res.values()))["nnz"]
    ncoef = cs["ncoef"]
    # compulsory bytes of a full assembly (mesh -> CSR/CSC + load vector)
    full_bytes = (E * (cs["nen"] + cs["nen_p"]) * 4 + Nn * 16 + E * 8 * ncoef + n * 4 + nnz * 12 + (n + 1) * 4
                  + (n * 8 if cs["heat"] else 0))
    # k_reuse with the gather fused in: conn + coordinates + coefficients in, values (+ F) out
    kernel_bytes = (E * (cs["nen"] + cs["nen_p"]) * 4 + Nn * 16 + E * 8 * ncoef + nnz * 8
                    + (n * 8 if cs["heat"] else 0))

    out = {}
    for key, r in res.items():
        ph = r["phases"]
        ms = r["wall_ms"]

        def per(name):
            return (ph[name][0] / ph[name][1]) if ph[name][1] else None
        out[key] = dict(ms_per_step=ms, event_ms_per_step=r["region_ms"], elements_per_s=E / (ms * 1e-3),
                        numeric_ms=per("numeric"), gather_ms=per("gather"), element_ms=per("element"),
                        sort_ms=per("sort"), reduce_ms=per("reduce"),
                        full_assembly_gbs=full_bytes / (ms * 1e-3) / 1e9, clock_ramp_steps=r["ramp"])
    errs = {}
    if "fp64" in res and "mixed" in res:
        a, b = res["mixed"], res["fp64"]
        assert np.array_equal(a["ptr"], b["ptr"]) and np.array_equal(a["ind"], b["ind"])
        errs["mixed_vs_fp64_frobenius"] = rel_frob(a["val"], b["val"])
        errs["mixed_vs_fp64_max_entry_over_max"] = float(np.abs(a["val"] - b["val"]).max() / np.abs(b["val"]).max())
        if cs["heat"]:
            errs["mixed_vs_fp64_load_vector"] = rel_frob(a["F"], b["F"])
    if "fp64" in res and "fp64_reuse" in res:
        errs["fp64_reuse_equals_fp64_plain"] = bool(np.array_equal(res["fp64"]["val"], res["fp64_reuse"]["val"]))

    m = out.get("mixed") or next(iter(out.values()))
    # the reuse step is the one fused kernel: its average duration is the
    # region events' time per step (per-launch events separate the launches
    # and read about 3 % higher: numeric_ms)
    kern_ms = m["event_ms_per_step"] if "mixed" in out or "fp64_reuse" in out else (m["numeric_ms"] or m["ms_per_step"])
    achieved = kernel_bytes / (kern_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": cs["kernel"], "kernel_ms": kern_ms, "kernel_ms_per_launch_events": m["numeric_ms"],
                "algorithmic_bytes": kernel_bytes}
    roofline.update(measured_traffic(args.config, args.N, cs["N"]))

    extra = {}
    if args.config == "c5" and not args.no_pcg:
        extra["pcg"] = bench_pcg(ctx, L, cs, res, args.pcg_maxit)
    if args.config == "c5":
        extra["device_setup"] = device_setup_c5(L, cs)
    if args.config == "c2":
        extra["element_matrices"] = bench_element_matrices(ctx, L, cs)
        progress("element matrices timed")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.config, cs)

    line = {
        "metric": METRIC,
        "value": m["elements_per_s"],
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 geometry, f32 element products, f64 sums",
        "data": "synthetic (seeded SplitMix64), inputs resident in HBM",
        "config": {"workload": cs["workload"], "elements": E, "nodes": Nn, "dofs": n, "nnz": nnz,
                   "layout": "CSR" if cs["layout"] == L.CSR else "CSC",
                   "method": "REUSE (pattern cached, fused owner-computes kernel)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "paths": out,
        "speedup_mixed_vs_fp64_plain": (out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"])
        if ("fp64" in out and "mixed" in out) else None,
        "full_assembly_bytes": full_bytes,
        "errors": errs,
        "geometry": {"min_det": mindet, "elements_over_quality_threshold": nbad},
        "setup": {"host_mesh_and_upload_s": t_setup, "symbolic_ms": t_sym},
        # the first mixed REUSE call on a new mesh and context: symbolic phase
        # (pattern, tiles, tables: on the device for elasticity and Stokes, host
        # C++ for heat) + the numeric kernel, allocations included, against the
        # plain FP64 method's first call (no symbolic phase) and its later calls
        "first_call": {"ms": t_sym, "elements_per_s": E / (t_sym * 1e-3),
                       "fp64_plain_first_call_ms": t_plain_first,
                       "fp64_plain_ms": out["fp64"]["ms_per_step"] if "fp64" in out else None,
                       "assemblies_to_break_even_vs_fp64_plain":
                           (t_sym / (out["fp64"]["ms_per_step"] - m["ms_per_step"]))
                           if "fp64" in out and out["fp64"]["ms_per_step"] > m["ms_per_step"] else None},
        "device": {"name": devname, "cus": ncu},
    }
    line.update(extra)
    print(json.dumps(with_copy_peak(line, ctx)))


RAMP_S = 0.3  # seconds of untimed back-to-back launches before the warm-up steps


def clock_ramp(launch, sync):
    """Untimed back-to-back launches for RAMP_S seconds before the warm-up
    steps: the GPU raises its clocks only under sustained load (C2 measured
    74.8 us per step after 20 warm-up steps, 62.5 us after 2000)."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < RAMP_S:
        for _ in range(16):
            launch()
        sync()
        n += 16
    return n


CPU_SAMPLE_S = 10.0  # seconds of single-thread CPU work in the timed CPU baseline sample

TRAFFIC_FILES = {"c1": "k_reuse_heat_t3_100x100_traffic.json",
                 "c2": "k_reuse_elasticity_t6_400x400_traffic.json",
                 "c3": "k_reuse_stokes_t6t3_500x500_traffic.json",
                 "c4": "k_reuse_batch_256_problems_traffic.json",
                 "c5": "k_reuse_heat_q4_4096x4096_and_pcg_traffic.json",
                 "t4": "k_fam_heat_t4_64cubed_traffic.json",
                 "p1b": "k_fam_reddy_p1b_p1_512x512_traffic.json"}


def kernel_source_hash():
    """SHA-256 of the library's device and host sources (csrc/*): a counter
    run recorded under one hash says nothing about the kernels of another."""
    import hashlib
    d = os.path.join(ROOT, "elfel.jl_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".h", ".cpp")) or name == "Makefile":
            with open(os.path.join(d, name), "rb") as fh:
                h.update(name.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def measured_traffic(cfgname, N_arg, N, block=None):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs of the same workload (scaled by the
    device-copy calibration), when the workload is the profiled size and the
    counters were taken on the current sources (same source hash); otherwise
    traffic stays null and the reason is given.  The batched kernel's file
    holds one entry per workgroup size."""
    f = TRAFFIC_FILES.get(cfgname)
    if not f or N_arg:
        return {}
    path = os.path.join(ROOT, "profiles", f)
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        t = json.load(fh)
    src = kernel_source_hash()
    if t.get("source_hash") != src:
        return {"traffic_note": f"counters in profiles/{f} were taken on other sources "
                                f"({t.get('source_hash')} against {src}); not used"}
    if "by_block" in t:
        t = t["by_block"].get(str(block))
        if not t:
            return {}
    return {"traffic": t["traffic_bytes"], "traffic_fetch": t["fetch_bytes"], "traffic_write": t["write_bytes"],
            "traffic_source": t["source"] + " (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, x2 / x1 copy calibration)",
            "source_hash": src}


def with_copy_peak(line, ctx):
    """The roofline beside a measured ceiling: a plain 1 GiB device copy
    (elfel_hip_copy_bandwidth, read + write bytes, 10 launches)."""
    try:
        peak = ctx.copy_bandwidth(1 << 30, 10)
    except Exception as e:  # the bench line stands without it
        line["roofline"]["copy_peak_note"] = f"copy probe failed: {e}"
        return line
    r = line["roofline"]
    r["copy_peak_gbs"] = round(peak, 1)
    r["frac_of_copy_peak"] = round(r["achieved"] / peak, 4)
    if r.get("traffic"):
        ms = r.get("kernel_ms") or r.get("kernel_ms_per_launch")
        if ms:
            r["traffic_gbs"] = round(r["traffic"] / (ms * 1e-3) / 1e9, 1)
    return line


def device_setup_c5(L, cs):
    """The heat-Q4 problem set up on the device instead of the host: block
    mesh, boundary selection and dof numbering (elfel_hip_block_mesh,
    elfel_hip_boundary_nodes, elfel_hip_number_dofs), checked against the
    host-built dof numbers; wall time of the three calls, host round trips
    included."""
    N = cs["N"]
    f = cs["cfg"]["fesp"]
    c2 = L.Context(0)
    c2.block_mesh(0, L.Q4, 1.0, 1.0, N, N)  # warm-up of the kernels and buffers
    c2.synchronize()
    t0 = time.perf_counter()
    E, Nn = c2.block_mesh(0, L.Q4, 1.0, 1.0, N, N)
    t1 = time.perf_counter()
    b = c2.boundary_nodes(0, Nn)
    t2 = time.perf_counter()
    isd = np.zeros((Nn, 1), bool)
    isd[b - 1] = True
    dn = c2.number_dofs([isd])[0]
    t3 = time.perf_counter()
    c2.close()
    return {"mesh_ms": (t1 - t0) * 1e3, "boundary_ms": (t2 - t1) * 1e3, "numbering_ms": (t3 - t2) * 1e3,
            "dofs_equal_host": bool(np.array_equal(dn, f.dofnums)), "elements": E, "nodes": Nn}


def bench_element_matrices(ctx, L, cs, reps=20):
    """The element matrices of the C2 mesh (T6 plane elasticity, 3-point rule,
    mixed) from the vector kernel and from the matrix-core kernel
    (v_mfma_f32_4x4x1f32), left on the device: median kernel time of `reps`
    calls each.  Both write 576 bytes per element (144 FP32 entries), so
    both are bound by the write of the blocks; the matrix-core form does the
    B^T D B contraction in 81 MFMA instructions per 16 elements."""
    E, D = cs["E"], cs["kw"]["D"]
    out = {"elements": E, "bytes_written": 576 * E}
    for name, meth in (("valu", L.EM_VALU), ("mfma", L.EM_MFMA)):
        ts = []
        for _ in range(reps + 3):
            ts.append(ctx.element_matrices(E, D, meth, copy=False)[1])
        ms = float(np.median(ts[3:]))
        out[f"{name}_ms"] = ms
        out[f"{name}_elements_per_s"] = E / (ms * 1e-3)
        out[f"{name}_write_gbs"] = 576 * E / (ms * 1e-3) / 1e9
    return out


def bench_pcg(ctx, L, cs, res, args_maxit=60000):
    """C5: Jacobi-PCG on the free block, FP32 inner iterations with FP64
    refinement against FP64 PCG, both to ||b - A x|| / ||b|| <= 1e-10."""
    cfg = cs["cfg"]
    nu, n = cfg["nfree"], cfg["ndofs"]
    # the matrix in the context must be the FP64 one
    ctx.assemble(cs["kind"], cs["fe"], cs["rule"], precision=L.FP64, method=L.REUSE, layout=cs["layout"],
                 **cs["kw"])
    F = ctx.copy_vector(n)
    x0 = np.zeros(n)
    outp = {}
    nnz_free = None
    for name, prec in (("mixed", L.MIXED), ("fp64", L.FP64)):
        ctx.synchronize()
        t0 = time.perf_counter()
        progress(f"pcg {name} ...")
        x, it, outer, rr = ctx.pcg(nu, F, x0, precision=prec, tol=1e-10, maxit=args_maxit)
        dt = time.perf_counter() - t0
        progress(f"pcg {name}: {it} iterations, {outer} residual updates, relres {rr:.2e}, {dt:.1f} s")
        outp[name] = dict(seconds=dt, iterations=it, refinements=outer, relres=rr, x=x)
    ptr = res["fp64"]["ptr"] if "fp64" in res else None
    if ptr is not None:
        nnz_free = int(ptr[nu])  # upper bound: rows < nu, all columns
    d = outp["mixed"]["x"][:nu] - outp["fp64"]["x"][:nu]
    sol_diff = float(np.linalg.norm(d) / max(np.linalg.norm(outp["fp64"]["x"][:nu]), 1e-300))
    for v in outp.values():
        v.pop("x")
    return {"unknowns": nu, "tol": 1e-10, "paths": outp, "mixed_vs_fp64_solution": sol_diff,
            "speedup_mixed_vs_fp64": outp["fp64"]["seconds"] / outp["mixed"]["seconds"],
            "nnz_rows_free": nnz_free}


# ---------------------------------------------------------------------------
# C4: batched uneven problems
# ---------------------------------------------------------------------------
def batch_fill(ctx, L, probs, precision):
    ctx.batch_begin()
    for p in probs:
        f = p["fesp"]
        if p["kind"] == "heat":
            ctx.batch_add(L.HEAT, L.T3, 1, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(),
                          coef_e=p["kappa_e"], kappa=1.0, Q=p["Q"], precision=precision)
        else:
            ctx.batch_add(L.ELAST, L.T6, 3, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(),
                          coef_e=p["scale_e"], D=p["D"], precision=precision)


def batch_time(ctx, steps, warmup):
    """(event ms per replay, wall ms per replay, clock-ramp launches)."""
    ctx.batch_assemble()
    ramp = clock_ramp(lambda: ctx.batch_assemble(wait=False), ctx.synchronize)
    for _ in range(warmup):
        ctx.batch_assemble(wait=False)
    ctx.synchronize()
    # one pair of hipEvents on the library's stream around the timed replays
    ctx.timing_region()
    w0 = time.perf_counter()
    for _ in range(steps):
        ctx.batch_assemble(wait=False)
    ctx.synchronize()
    wall = (time.perf_counter() - w0) * 1e3 / steps
    ph = ctx.timing_read()
    return ph["total"][0] / max(1, ph["total"][1]), wall, ramp


def bench_batch(args, L, W):
    ctx = L.Context(0)
    devname, ncu = ctx.device_name()
    specs = W.batch_specs()
    t0 = time.time()
    probs = [W.batch_problem(k, kind, nn) for k, (kind, nn) in enumerate(specs)
             if args.kinds in ("all", kind)]
    t_host = time.time() - t0
    E = Nn = conn_b = coef_b = F_b = 0
    for p in probs:
        f = p["fesp"]
        if p["kind"] == "heat":
            F_b += f.ndofs() * 8
        E += f.mesh.nelements
        Nn += f.mesh.nnodes
        conn_b += f.mesh.conn.size * 4
        coef_b += f.mesh.nelements * 8
    batch_fill(ctx, L, probs, L.MIXED)
    t0 = time.perf_counter()
    ctx.batch_build(autotune=True)
    t_build = (time.perf_counter() - t0) * 1e3
    st = ctx.batch_stats()
    ms, wall, ramp = batch_time(ctx, args.steps, args.warmup)
    progress("mixed batch timed")
    info0 = ctx.batch_info(0)
    # the same batch in FP64 (the FP64 REUSE arithmetic in the same launch)
    c64 = L.Context(0)
    batch_fill(c64, L, probs, L.FP64)
    c64.batch_build(autotune=True)
    ms64, wall64, _ = batch_time(c64, max(1, args.steps // 4), min(args.warmup, 5))
    progress("fp64 batch timed")
    err64 = 0.0
    for k in range(len(probs)):
        a, b = ctx.batch_copy(k), c64.batch_copy(k)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        err64 = max(err64, rel_frob(a[2], b[2]))
    c64.close()
    kernel_bytes = conn_b + Nn * 16 + coef_b + st["nnz"] * 8 + F_b
    # errors of a sample of problems against the FP64 plain method
    errs = []
    fp64_ms = 0.0
    sample = list(range(0, len(probs), max(1, len(probs) // 16)))
    ctx2 = L.Context(0)
    for k in sample:
        p = probs[k]
        f = p["fesp"]
        ctx2.set_mesh(0, f.mesh.conn, f.mesh.xy)
        ctx2.set_dofs(0, f.dofnums, f.ndofs())
        if p["kind"] == "heat":
            ctx2.set_coefficients(f.mesh.nelements, p["kappa_e"], None, None)
            a = dict(kind=L.HEAT, fe=L.T3, rule=1, kw=dict(kappa=1.0, Q=p["Q"]))
        else:
            ctx2.set_coefficients(f.mesh.nelements, None, None, p["scale_e"])
            a = dict(kind=L.ELAST, fe=L.T6, rule=3, kw=dict(D=p["D"]))
        ctx2.assemble(a["kind"], a["fe"], a["rule"], precision=L.FP64, method=L.SORT, layout=L.CSC, **a["kw"])
        ctx2.synchronize()
        ctx2.timing(True)
        ctx2.assemble(a["kind"], a["fe"], a["rule"], precision=L.FP64, method=L.SORT, layout=L.CSC, **a["kw"])
        ctx2.synchronize()
        fp64_ms += ctx2.timing_read()["total"][0]
        ctx2.timing(False)
        ref = ctx2.copy_sparse(base=0)
        got = ctx.batch_copy(k)
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
        errs.append(rel_frob(got[2], ref[2]))
    E_sample = sum(probs[k]["fesp"].mesh.nelements for k in sample)
    fp64_eps = E_sample / (fp64_ms * 1e-3)
    line = {
        "metric": METRIC, "value": E / (ms * 1e-3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 geometry, f32 element products, f64 sums",
        "data": "synthetic (seeded SplitMix64), inputs resident in HBM",
        "config": {"workload": ("256 uneven heat-T3 (1-pt) / elasticity-T6 (3-pt) problems, E_k log-uniform "
                                "[1e3, 2e5], lognormal(0, 0.5) coefficients, +-20% h jitter, one persistent "
                                "launch replayed from a hipGraph")
                   + ("" if args.kinds == "all" else f"; only its {len(probs)} {args.kinds} problems"),
                   "problems": len(probs), "elements": E, "nnz": st["nnz"], "tiles": st["ntiles"],
                   "block": info0["block"], "grid": info0["grid"]},
        "roofline": {"bound": "hbm", "achieved": round(kernel_bytes / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(kernel_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_reuse_batch (hipGraph replay)", "algorithmic_bytes": kernel_bytes},
        "cpu_baseline": None,
        "wall_ms_per_step": wall,
        "clock_ramp_steps": ramp,
        "fp64_batch": {"ms_per_step": ms64, "wall_ms_per_step": wall64, "elements_per_s": E / (ms64 * 1e-3)},
        "speedup_mixed_vs_fp64_batch": ms64 / ms,
        "fp64_plain_elements_per_s_sample": fp64_eps,
        "speedup_mixed_vs_fp64_plain": (E / (ms * 1e-3)) / fp64_eps,
        "errors": {"max_frobenius_vs_fp64_sample": max(errs), "problems_checked": len(sample),
                   "max_frobenius_vs_fp64_batch_all_problems": err64},
        "setup": {"host_meshes_s": t_host, "build_and_autotune_ms": t_build, "tuned_ms": st["tuned_ms"]},
        "device": {"name": devname, "cus": ncu},
    }
    if args.kinds == "all":
        line["roofline"].update(measured_traffic("c4", 0, 0, block=info0["block"]))
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_batch(probs, sample)
    print(json.dumps(with_copy_peak(line, ctx)))


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


# ---------------------------------------------------------------------------
# CPU baseline: the oracle (C, one thread), element loop + sparse()
# ---------------------------------------------------------------------------
def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def _oracle_problem(O, o):
    f = o["fesp"]
    fe = {"T3": O.T3, "T6": O.T6, "Q4": O.Q4}[f.fe.name]
    if o["kind"] == "heat":
        rule = 2 if f.fe.name == "Q4" else 1
        return O.make_problem(O.HEAT, fe, rule, f.mesh.conn, f.mesh.xy, f.dofnums, kappa=1.0,
                              Q=o.get("Q", -6.0), kappa_e=o.get("kappa_e"), Q_e=o.get("Q_e"))
    if o["kind"] == "elast":
        return O.make_problem(O.ELAST, fe, 3, f.mesh.conn, f.mesh.xy, f.dofnums, D=o["D"],
                              scale_e=o.get("scale_e"))
    p = o["pfesp"]
    return O.make_problem(O.STOKES_GEN, fe, 3, f.mesh.conn, f.mesh.xy, f.dofnums, conn_p=p.mesh.conn,
                          dofnums_p=p.dofnums, D=o["D"], mu=1.0)


def cpu_baseline(cfgname, cs):
    try:
        O = _oracle()
    except Exception as e:  # the checker is optional for the GPU numbers
        return {"error": f"oracle unavailable: {e}"}
    from elfel_hip import workloads as W
    sample_cs = cs
    sample = None
    if cfgname == "c5" and cs["N"] > 1024:
        cfg = W.heat_q4(1024, sigma=1.0)
        f = cfg["fesp"]
        sample_cs = dict(oracle=dict(kind="heat", fesp=f, kappa_e=cfg["kappa_e"], Q_e=cfg["Q_e"]),
                         E=f.mesh.nelements)
        sample = "heat Q4 1024x1024 (1,048,576 elements) of the same construction"
    if cfgname == "c3" and cs["N"] > 250:
        cfg = W.stokes_t6t3(250)
        sample_cs = dict(oracle=dict(kind="stokes", fesp=cfg["ufesp"], pfesp=cfg["pfesp"], D=cfg["D"]),
                         E=cfg["ufesp"].mesh.nelements)
        sample = "Stokes T6/T3 250x250 (125,000 elements) of the same construction"
    p = _oracle_problem(O, sample_cs["oracle"])
    te, ts, nnz = O.time_assembly(p, reps=1, nthreads=1)
    if te + ts < 0.1:  # a cold first pass misjudges small samples
        te, ts, nnz = O.time_assembly(p, reps=max(1, int(0.2 / max(te + ts, 1e-6))), nthreads=1)
    # repeat the sample for about CPU_SAMPLE_S seconds of CPU work
    reps = max(1, min(5000, int(round(CPU_SAMPLE_S / max(te + ts, 1e-6)))))
    if reps > 1:
        te, ts, nnz = O.time_assembly(p, reps=reps, nthreads=1)
    E = sample_cs["E"]
    # the same element loop on every host core (sparse() stays serial), for reference
    nth = max(1, min(16, os.cpu_count() or 1))
    te_n, ts_n, _ = O.time_assembly(p, reps=reps, nthreads=nth)
    return {"value": E / (te + ts), "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": (sample or f"the full mesh ({E} elements)") + f", {reps} repetitions ({reps * (te + ts):.1f} s): "
                      f"element loop {te:.4f} s + sparse() {ts:.4f} s each, single thread (the reference's path "
                      f"is serial)",
            "element_loop_s": te, "sparse_s": ts, "repetitions": reps, "host_cores_visible": os.cpu_count(),
            "threaded_element_loop": {"threads": nth, "element_loop_s": te_n, "sparse_s": ts_n,
                                      "elements_per_s": E / (te_n + ts_n)}}


def cpu_baseline_batch(probs, sample):
    try:
        O = _oracle()
    except Exception as e:
        return {"error": f"oracle unavailable: {e}"}
    ops = []
    E = 0
    for k in sample:
        p = probs[k]
        o = dict(kind=p["kind"], fesp=p["fesp"], kappa_e=p.get("kappa_e"), Q=p.get("Q", 0.0),
                 D=p.get("D"), scale_e=p.get("scale_e"))
        ops.append(_oracle_problem(O, o))
        E += p["fesp"].mesh.nelements

    def run(reps):
        te = ts = 0.0
        for q in ops:
            a, b, _ = O.time_assembly(q, reps=reps, nthreads=1)
            te, ts = te + a, ts + b
        return te, ts
    te, ts = run(1)
    reps = max(1, min(5000, int(round(CPU_SAMPLE_S / max(te + ts, 1e-6)))))
    if reps > 1:
        te, ts = run(reps)
    return {"value": E / (te + ts), "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": f"{len(sample)} of the 256 problems ({E} elements), {reps} repetitions "
                      f"({reps * (te + ts):.1f} s): element loop {te:.4f} s + sparse() {ts:.4f} s each, single thread"}


# ---------------------------------------------------------------------------
# element families (REUSE method with a cached sort permutation; plain FP64)
# ---------------------------------------------------------------------------
def family_case(cfgname, N):
    """Mesh and numbering of a family workload (host arrays)."""
    from elfel_hip import meshes as M
    if cfgname == "t4":
        N = N or 64
        m = M.T4block(1.0, 1.0, 1.0, N, N, N)
        E, Nn = m.nelements, m.nnodes
        return dict(N=N, E=E, Nn=Nn, n=Nn, conn=m.conn, x=m.xy, dn=np.arange(1, Nn + 1),
                    workload=f"heat T4 {N}^3 block (six tetrahedra per cell), 1-pt rule, kappa 1, Q -6, CSR + "
                             f"load vector",
                    kernel="k_fam_scalar<HEAT, 3D, T4, 1 point>",
                    kernel_in=E * 16 + Nn * 24 + Nn * 4, per_el_out=16 * 16 + 4 * 16,
                    mesh_bytes=E * 16 + Nn * 24 + Nn * 4, heat=True,
                    geom_in=E * 16 + Nn * 24, entries=16, fentries=4)
    N = N or 512
    m = M.transform(M.T3block(2.0, 2.0, N, N), lambda x: x - 1.0)
    E, Nn = m.nelements, m.nnodes
    dux, duy = np.arange(1, Nn + 1), np.arange(Nn + 1, 2 * Nn + 1)
    eux, euy = np.arange(2 * Nn + 1, 2 * Nn + E + 1), np.arange(2 * Nn + E + 1, 2 * Nn + 2 * E + 1)
    dp = np.arange(2 * Nn + 2 * E + 1, 3 * Nn + 2 * E + 1)
    return dict(N=N, E=E, Nn=Nn, n=3 * Nn + 2 * E, conn=m.conn, x=m.xy, dux=dux, duy=duy, eux=eux, euy=euy, dp=dp,
                workload=f"Reddy Stokes P1b/P1 {N}x{N} (T3 + bubble velocity, T3 pressure), 3-pt rule, CSR",
                kernel="k_fam_reddy<T3B, T3, 3 points>",
                kernel_in=E * 24 + Nn * 16 + Nn * 12 + E * 8, per_el_out=(4 * 16 + 4 * 12) * 16,
                mesh_bytes=E * 24 + Nn * 16 + Nn * 12 + E * 8, heat=False,
                geom_in=E * 12 + Nn * 16, entries=4 * 16 + 4 * 12, fentries=0)


def family_upload(ctx, L, cs):
    """Uploads the workload; returns assemble(method, precision, wait)."""
    if "dn" in cs:
        ctx.set_mesh(0, cs["conn"], cs["x"])
        ctx.set_dofs(0, cs["dn"], cs["n"])
        return lambda m=L.SORT, p=L.FP64, wait=True: ctx.assemble(L.HEAT, L.T4, 1, precision=p, method=m,
                                                                  layout=L.CSR, kappa=1.0, Q=-6.0, wait=wait)
    ctx.set_mesh(0, cs["conn"], cs["x"])
    ctx.set_dofs(0, np.stack([cs["dux"], cs["duy"]], 1), cs["n"])
    ctx.set_element_dofs(0, np.stack([cs["eux"], cs["euy"]], 1), cs["n"])
    ctx.set_mesh(1, cs["conn"], cs["x"])
    ctx.set_dofs(1, cs["dp"], cs["n"])
    return lambda m=L.SORT, p=L.FP64, wait=True: ctx.assemble(L.STOKES_REDDY, L.T3B, 3, precision=p, method=m,
                                                              layout=L.CSR, mu=1.0, fe_p=L.T3, wait=wait)


def family_oracle(O, cs):
    if "dn" in cs:
        I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT, O.T4, 1, cs["conn"], cs["x"], cs["dn"], kappa=1.0, Q=-6.0)
        O.sparse(I, J, V, cs["n"], cs["n"], csc=False)
        O.vec_assemble(FI, FV, cs["n"])
        return
    I, J, V = O.fam_reddy_coo(O.T3B, O.T3, 3, cs["conn"], cs["x"], cs["dux"], cs["duy"], mu=1.0, de_ux=cs["eux"],
                              de_uy=cs["euy"], conn_p=cs["conn"], xy_p=cs["x"], dn_p=cs["dp"])
    O.sparse(I, J, V, cs["n"], cs["n"], csc=False)


def bench_family(args, L):
    ctx = L.Context(0)
    devname, ncu = ctx.device_name()
    t0 = time.time()
    cs = family_case(args.config, args.N)
    assemble = family_upload(ctx, L, cs)
    t_setup = time.time() - t0
    progress(f"{args.config}: mesh, numbering and upload ready ({t_setup:.1f} s)")
    E, Nn, n = cs["E"], cs["Nn"], cs["n"]

    def timed(m, p, steps):
        for _ in range(args.warmup):
            assemble(m, p, False)
        ctx.synchronize()
        w0 = time.perf_counter()
        for _ in range(steps):
            assemble(m, p, False)
        ctx.synchronize()
        ms = (time.perf_counter() - w0) * 1e3 / steps
        ctx.timing(True)
        for _ in range(steps):
            assemble(m, p, False)
        ctx.synchronize()
        ph = ctx.timing_read()
        ctx.timing(False)
        return ms, {k: (ph[k][0] / ph[k][1]) if ph[k][1] else None for k in ("element", "sort", "reduce", "total")}

    # the plain FP64 method, every step
    nnz = assemble()
    ctx.synchronize()
    ptr0, ind0, val0 = ctx.copy_sparse(base=0)[:3]
    ramp = clock_ramp(lambda: assemble(L.SORT, L.FP64, False), ctx.synchronize)
    ms_plain, ph_plain = timed(L.SORT, L.FP64, min(args.steps, 50))
    progress("fp64 plain timed")
    # first REUSE call on this mesh: the symbolic phase (one plain-method pass
    # that sorts stream positions) and the numeric phase
    ctx.synchronize()
    w0 = time.perf_counter()
    assemble(L.REUSE, L.MIXED)
    first_ms = (time.perf_counter() - w0) * 1e3
    ms_mixed, ph_mixed = timed(L.REUSE, L.MIXED, args.steps)
    ptr, ind, valm = ctx.copy_sparse(base=0)[:3]
    same_pattern = bool(np.array_equal(ptr, ptr0) and np.array_equal(ind, ind0))
    err = float(np.linalg.norm(valm - val0) / max(np.linalg.norm(val0), 1e-300))
    ms_r64, ph_r64 = timed(L.REUSE, L.FP64, args.steps)
    val64 = ctx.copy_sparse(base=0)[2]
    reuse_bitwise = bool(np.array_equal(val64, val0))
    progress("reuse timed")
    # the mixed step's two kernels: element values (geometry in, FP32 values
    # out in stream order) and the runs (permutation, run starts and the
    # values in, FP64 matrix and load vector out)
    nent, nf = E * cs["entries"], E * cs["fentries"]
    el_bytes = cs["geom_in"] + 4 * (nent + nf)
    runs_bytes = 4 * nent + 4 * (nnz + 1) + 4 * nent + 8 * nnz
    if cs["heat"]:
        runs_bytes += 4 * nf + 4 * (n + 1) + 4 * nf + 8 * n
    el_ms, runs_ms = ph_mixed["element"], ph_mixed["reduce"]
    dom = "element" if el_ms >= runs_ms else "runs"
    kernel_bytes, kernel_ms = (el_bytes, el_ms) if dom == "element" else (runs_bytes, runs_ms)
    kname = (cs["kernel"] + " (values only, FP32 out)") if dom == "element" else "k_fam_runs<float>"
    achieved = kernel_bytes / (kernel_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        try:
            O = _oracle()
            scs = family_case(args.config, 32 if args.config == "t4" else 128)
            t1 = time.perf_counter()
            family_oracle(O, scs)
            one = time.perf_counter() - t1
            reps = max(1, min(1000, int(round(CPU_SAMPLE_S / max(one, 1e-6)))))
            t1 = time.perf_counter()
            for _ in range(reps):
                family_oracle(O, scs)
            each = (time.perf_counter() - t1) / reps
            cpu = {"value": scs["E"] / each, "unit": "elements/s", "cores": 1, "kind": "port",
                   "sample": f"{scs['workload']} ({scs['E']} elements), {reps} repetitions ({reps * each:.1f} s): "
                             f"element loop + sparse() {each:.4f} s each, single thread"}
        except Exception as e:  # the checker is optional for the GPU numbers
            cpu = {"error": f"oracle unavailable: {e}"}
    line = {
        "metric": METRIC,
        "value": E / (ms_mixed * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 geometry, f32 element products, f64 sums",
        "data": "synthetic structured mesh, inputs resident in HBM",
        "config": {"workload": cs["workload"], "elements": E, "nodes": Nn, "dofs": n, "nnz": nnz, "layout": "CSR",
                   "method": "REUSE: element kernel (values only) -> runs summed through the cached sort permutation"},
        "roofline": dict({"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kname,
                          "kernel_ms": kernel_ms, "algorithmic_bytes": kernel_bytes},
                         **measured_traffic(args.config, args.N, cs["N"])),
        "cpu_baseline": cpu,
        "phases_ms": ph_mixed,
        "kernels": {"element_values": {"ms": el_ms, "bytes": el_bytes, "gbs": el_bytes / (el_ms * 1e-3) / 1e9},
                    "runs": {"ms": runs_ms, "bytes": runs_bytes, "gbs": runs_bytes / (runs_ms * 1e-3) / 1e9}},
        "mixed_vs_fp64_frobenius": err,
        "same_pattern_as_plain": same_pattern,
        "fp64_reuse": {"ms_per_step": ms_r64, "elements_per_s": E / (ms_r64 * 1e-3), "phases_ms": ph_r64,
                       "equals_fp64_plain_bitwise": reuse_bitwise},
        "fp64_plain": {"ms_per_step": ms_plain, "elements_per_s": E / (ms_plain * 1e-3), "phases_ms": ph_plain,
                       "kernel": cs["kernel"]},
        "speedup_mixed_vs_fp64_plain": ms_plain / ms_mixed,
        "speedup_mixed_vs_fp64_reuse": ms_r64 / ms_mixed,
        "first_call": {"ms": first_ms, "elements_per_s": E / (first_ms * 1e-3),
                       "fp64_plain_ms": ms_plain},
        "clock_ramp_steps": ramp,
        "setup": {"host_mesh_numbering_upload_s": t_setup},
        "device": {"name": devname, "cus": ncu},
    }
    print(json.dumps(with_copy_peak(line, ctx)))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU per box; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    from elfel_hip import _lib as L
    from elfel_hip import workloads as W
    if args.config == "c4":
        bench_batch(args, L, W)
    elif args.config in ("t4", "p1b"):
        bench_family(args, L)
    else:
        bench_single(args, L, W)


if __name__ == "__main__":
    main()
