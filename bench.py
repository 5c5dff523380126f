#!/usr/bin/env python3
"""SPR size-field benchmark on one MI355X.

Workload (strain_120): the unit cube as 120^3 hexes x 6 Kuhn tets
(10,368,000 tets, 1,771,561 vertices, 12,225,960 edges), Lagrange-2
coordinates, interior jitter 0.25h (seed 0x5EED0005); a 9-component strain
IP field eps = sym(grad v), v = (sin(pi x) y, z^2, cos(pi y) x), at the 4
points of the order-2 tet rule; one step = spr::getSPRSizeField(eps, 0.05)
(patch recovery of all vertex and edge patches, element error integration,
size factor, vertex average) with the IP field resident in HBM.

Prints one JSON line.  value = recovered patches (V + E) per second of the
whole size-field step in the chosen form (default fp64f: FP64 normal
equations, the fastest form that meets 1e-6 relative on the size field);
"forms" reports the other two forms on the same mesh (FP32 + FP64 refinement,
and the exact FP64 form that reproduces the reference bit for bit) with
their accuracy against the exact form; roofline = the dominant kernel (the
patch fit over every patch, one launch) with the recovery's algorithmic HBM
bytes over its GPU time (HIP events the library records on its stream);
cpu_baseline = the FP64 CPU oracle's whole step on the host cores, RCB-
partitioned with a two-layer ghost halo (one forked process per part), plus
one part's share on one core; its recovered field is checked bit for bit
against the GPU's FP64 form.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(T, V, E, q=4, c=9, order=2):
    """Compulsory HBM bytes per phase (FP64 data, int32 indices)."""
    N = V + (E if order == 2 else 0)
    k = 10 if order == 2 else 4
    up = 4 * T + (6 * T if order == 2 else 0)
    recover = T * q * c * 8 + N * 24 + T * k * 4 + up * 4 + (N + 1) * 4 + N * c * 8
    size = T * q * c * 8 + N * c * 8 + T * k * 4 + N * 24 + T * 8
    average = T * 8 + 4 * T * 4 + (V + 1) * 4 + V * 8
    return {"recover": recover, "size": size, "average": average}


def fit_flops(T, V, E, q=4, order=2):
    """Algorithmic FP64 flops of the fast form's regular fit: every patch
    fits q rows per tet of its set (each tet is in 4 vertex and, on a P2
    mesh, 6 edge patches); per row 74 flops for the Gram moments (3 sub,
    6 mul, 3 add, 31 FMA) and 56 for the weighted sum (3 sub, 6 mul, 10 FMA,
    9 sub, 9 FMA of 9 components); per patch about 900 for the 10x10
    Cholesky factor and its two triangular solves."""
    rows = T * q * (4 + (6 if order == 2 else 0))
    return 130 * rows + 900 * (V + (E if order == 2 else 0))


def copy_probe(torch, nbytes=1 << 30, reps=5):
    """Device-to-device copy of nbytes (a plain read+write kernel): the
    measured HBM rate (2 * nbytes per copy) and the reference for scaling
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings."""
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del src, dst
    return {"bytes_per_copy": 2 * nbytes, "best_ms": best, "gbs": 2 * nbytes / (best / 1e3) / 1e9}


def cpu_baseline(host, ip, workers):
    """The FP64 oracle's whole step, partitioned: RCB parts with a two-layer
    ghost halo, one forked process per part (oracle/partition.py), timed
    start to end; then part 0's share alone on one core.  Returns the
    record and the CPU run's size and recovered fields (checked against the
    GPU's FP64 form later).  Call it before the GPU is touched: the parts
    are forked."""
    from oracle import partition

    nproc, avail = partition.cores()
    size, rec, info = partition.spr_size(host, ip, 2, 0.05, workers, serial_part=0)
    V, E, T = host.nv, host.ne, host.nt
    wall = info["wall_s"]
    ph = info["phase_s"]
    c = info["serial_counts"]
    sp = c["vertices"] + c["edges"]
    out = {
        "value": (V + E) / wall,
        "unit": "patches/s",
        "cores": workers,
        "kind": "port",
        "nproc": nproc,
        "cores_available": avail,
        "step_s": wall,
        "sample": ("the whole step, not a sample: getSPRSizeField(0.05) over all %d patches and %d "
                   "tets with the FP64 oracle (reference operation order, explicit-Q Householder), "
                   "RCB-partitioned into %d parts with a 2-layer ghost halo, one forked process per "
                   "part; host nproc %d, %d cores available, %d used; %.2f s start to end "
                   "(per part, barrier waits excluded: recovery %.2f-%.2f s, errors up to %.2f s, "
                   "sizes up to %.3f s, average up to %.3f s)"
                   % (V + E, T, workers, nproc, avail, workers, wall, ph[:, 0].min(), ph[:, 0].max(),
                      ph[:, 1].max(), ph[:, 2].max(), ph[:, 3].max())),
        "serial": {
            "value": sp / info["serial_s"], "unit": "patches/s", "cores": 1, "kind": "port",
            "sample": ("part 0 of the %d alone on one core, after the partitioned step: the step's "
                       "work for its %d vertex + %d edge patches and %d tets (local mesh %d tets), "
                       "%.2f s" % (workers, c["vertices"], c["edges"], c["tets"], c["local_tets"],
                                   info["serial_s"]))},
    }
    return out, size, rec


def grad_bytes(T, N, c_in=3, q=4, k=10):
    """Compulsory HBM bytes of getGradIPField: nodal values and coordinates
    of the N nodes, the T tets' k node ids, the T x q x 3 c_in gradients."""
    return N * c_in * 8 + N * 24 + T * k * 4 + T * q * 3 * c_in * 8


def grad_main(args, np, torch, spr, workloads):
    """--op grad: spr::getGradIPField of a Lagrange-2 vector field (3
    components -> 9 values at each of the 4 points of the order-2 rule) on
    the same mesh; one step = one call with the nodal field resident in HBM;
    value = tets per second of the call; roofline = the gradient kernel's
    algorithmic bytes over its GPU time (HIP events the library records on
    its stream); cpu_baseline = the FP64 oracle's gradient of the mesh's
    first 2^20 tets on one core."""
    t0 = time.time()
    host = workloads.strain_mesh(args.n)
    P = workloads.node_points(host)
    u = np.stack([workloads.sin_field(P), workloads.exp_field(P), P[:, 0] * P[:, 1]], 1)
    log("mesh n=%d: %d tets, %d nodes (%.1fs)" % (args.n, host.nt, P.shape[0], time.time() - t0))
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        w = min(host.nt, 1 << 20)
        sub = O.Mesh(host.tv[:w], host.xyz, host.order, te=host.te[:w], ev=host.ev)
        t1 = time.perf_counter()
        O.grad_ip(sub, 2, u, 2)
        dt = time.perf_counter() - t1
        cpu = {"value": w / dt, "unit": "tets/s", "cores": 1, "kind": "port",
               "sample": "the FP64 oracle's getGradIPField over the mesh's first %d tets "
                         "(4 points, 9 values each) on one core, %.2f s" % (w, dt)}
        log("cpu baseline: %s" % cpu)
    probe = copy_probe(torch)
    m = spr.Mesh(host)
    fu = m.createLagrangeField("u", 3, 2, u)
    torch.cuda.synchronize()

    def step():
        g = spr.getGradIPField(fu, "g", 2)  # returns after the kernel is complete
        m.destroyField(g)
        return m.phase_times()[0]

    for i in range(args.warmup):
        step()
    kernel_ms = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        kernel_ms.append(step())
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t_start) * 1e3 / args.steps
    kms = float(np.mean(kernel_ms))
    T, N = host.nt, P.shape[0]
    gb = grad_bytes(T, N)
    achieved = gb / (kms / 1e3) / 1e9
    out = {
        "metric": "getGradIPField tets/s @10.4M tets (P2 vector field, IP order 2)",
        "value": T / (ms_per_step / 1e3),
        "unit": "tets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "strain_120 mesh: getGradIPField of a P2 vector field at IP order 2, "
                               "10,368,000 tets, 3 -> 9 components", "tets": T, "nodes": N},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "measured_copy_gbs": probe["gbs"],
                     "kernel": "k_grad, %d algorithmic bytes per launch, %.3f ms" % (gb, kms)},
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


def traffic_of(form):
    """HBM bytes per launch of the fit pass from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement (scaled by the copy probe), for the
    form it was measured on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "k_tile_strain120_traffic.json")
    if form != "fp64f" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["traffic_bytes"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", default="fp64f", choices=["fp64", "fp32r", "fp64f"])
    ap.add_argument("--n", type=int, default=120)
    ap.add_argument("--cpu-workers", type=int, default=None,
                    help="processes of the partitioned CPU baseline (default: the cores available, "
                         "at most 16, the host share of one GPU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--exact-steps", type=int, default=2)
    ap.add_argument("--autotune", action="store_true",
                    help="choose the patch kernels' launch shapes by timing before the timed steps")
    ap.add_argument("--tune-cache", default=None,
                    help="JSON cache of tuned launch shapes (read, and written after tuning)")
    ap.add_argument("--tile-blocks", type=int, default=None,
                    help="grid cap of the tiled regular fit pass (0: thread per patch instead)")
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="run the 1 GiB device copy probe and exit")
    ap.add_argument("--op", default="size", choices=["size", "grad"],
                    help="size: getSPRSizeField (the default measurement); grad: getGradIPField")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; runs on several GPUs are not supported "
              "(each GPU would take one mesh part)", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from core_amd import spr, workloads

    if args.copy_probe_only:
        print(json.dumps({"copy_probe": copy_probe(torch)}), flush=True)
        return 0
    if args.op == "grad":
        return grad_main(args, np, torch, spr, workloads)

    t0 = time.time()
    host = workloads.strain_mesh(args.n)
    log("mesh n=%d: %d tets, %d vertices, %d edges (%.1fs)" % (args.n, host.nt, host.nv, host.ne,
                                                                time.time() - t0))
    ip = workloads.strain_ip(host, progress=lambda d, n: log("  strain IP field %d/%d tets" % (d, n))
                             if d % (1 << 23) < (1 << 20) or d == n else None)
    log("strain IP field %s (%.1fs)" % (ip.shape, time.time() - t0))
    cpu = cpu_size = cpu_rec = None
    if not args.no_cpu:
        # before the GPU is initialised: the partitioned baseline forks
        from oracle import partition

        workers = args.cpu_workers or min(16, partition.cores()[1])
        log("cpu baseline: whole step over %d parts" % workers)
        t1 = time.time()
        cpu, cpu_size, cpu_rec = cpu_baseline(host, ip, workers)
        log("cpu baseline (%.1fs): %s" % (time.time() - t1, cpu))
    probe = copy_probe(torch)
    log("copy probe: %s" % probe)
    m = spr.Mesh(host)
    eps = m.createIPField("eps", 9, 2, ip)
    torch.cuda.synchronize()
    log("uploaded (%.1fs)" % (time.time() - t0))
    shape = None
    if args.tile_blocks is not None:
        shape = {"fast_threads": 256, "fast_blocks": 8192, "exact_blocks": 8192,
                 "expand_blocks": 4096, "tile_blocks": args.tile_blocks}
        m.set_shape(shape)
    if args.autotune:
        t1 = time.time()
        shape = m.autotune(eps, args.form, cache=args.tune_cache)
        log("autotuned launch shapes %s (%.1fs)" % (shape, time.time() - t1))

    def step(form):
        s = spr.getSPRSizeField(eps, 0.05, form=form)
        m.destroyField(s)
        return m.phase_times()

    per_step = {}

    def timed_steps(form, warmup, steps):
        for i in range(warmup):
            step(form)
        phases = np.zeros(5)
        each = []
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for i in range(steps):
            t = time.perf_counter()
            phases += step(form)  # returns after the call's work is complete
            each.append((time.perf_counter() - t) * 1e3)
        torch.cuda.synchronize()
        t_end = time.perf_counter()
        per_step[form] = each
        return (t_end - t_start) * 1e3 / steps, phases / steps

    ms_per_step, phases = timed_steps(args.form, args.warmup, args.steps)
    routes = m.last_stats()
    log("%s: timed %d steps: %.3f ms/step; phases ms %s" % (
        args.form, args.steps, ms_per_step, np.array2string(phases, precision=3)))
    T, V, E = host.nt, host.nv, host.ne
    patches = V + E
    value = patches / (ms_per_step / 1e3)
    ab = algorithmic_bytes(T, V, E)
    # dominant kernel: the fit of every patch (one launch; phase 0).  In the
    # exact form phase 0 is the classify pass plus the bucketed fits.
    fit_ms = float(phases[0])
    achieved = ab["recover"] / (fit_ms / 1e3) / 1e9
    out = {
        "metric": "SPR patches/s & size-field ms @10.4M tets: FP32+refine vs FP64, rel err<=1e-6",
        "value": value,
        "unit": "patches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "ms_per_step_median": float(np.median(per_step[args.form])),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.form == "fp32r" else "f64",
        "data": "synthetic",
        "config": {"workload": "strain_120: getSPRSizeField(0.05) on n=120 P2 jittered cube, "
                               "10,368,000 tets, 9-comp strain IP field, IP order 2",
                   "form": args.form, "tets": T, "vertices": V, "edges": E, "patches": patches,
                   "launch_shape": shape or "default"},
        "phases_ms": {"fit": phases[0], "expand": phases[1], "expanded_fit": phases[2],
                      "element": phases[3], "factor_average": phases[4]},
        "routes": routes,
        "roofline_fp64": ({"bound": "fp64 vector ALU", "achieved": fit_flops(T, V, E) / (fit_ms / 1e3) / 1e12,
                           "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": fit_flops(T, V, E) / (fit_ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS,
                           "flops": fit_flops(T, V, E),
                           "note": "the same launch priced by its algorithmic FP64 flops, the bound it "
                                   "is held by (its HBM traffic moves at a fraction of peak)"}
                          if args.form == "fp64f" else None),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(args.form),
                     "measured_copy_gbs": probe["gbs"],
                     "kernel": ("k_tile regular pass" if args.form != "fp64" else "exact fit")
                               + " (all %d patches), %d algorithmic bytes per launch, %.3f ms"
                               % (patches, ab["recover"], fit_ms)},
    }
    if not args.no_compare:
        # the other forms on the same mesh; accuracy against the exact FP64
        # form (bit-identical to the FP64 oracle)
        s64 = spr.getSPRSizeField(eps, 0.05, form="fp64")
        a64 = s64.data[:, 0].clone()
        m.destroyField(s64)
        r64 = spr.recoverField(eps, form="fp64")
        d64 = r64.data.clone()
        m.destroyField(r64)
        if cpu is not None:
            # the partitioned CPU run against the GPU's FP64 form: the same
            # recovered field bit for bit, the size field to summation order
            same = bool(torch.equal(torch.from_numpy(cpu_rec).to(d64.device), d64))
            srel = float((torch.from_numpy(cpu_size).to(a64.device) / a64 - 1).abs().max())
            cpu["agrees_with_gpu_fp64"] = {"recovered_bitwise": same, "size_max_rel": srel}
            log("cpu baseline vs GPU fp64: %s" % cpu["agrees_with_gpu_fp64"])
            cpu_rec = cpu_size = None
        scale = d64.abs().amax(0)
        keep = scale > 0
        forms = {}
        for form in ["fp64f", "fp32r", "fp64"]:
            if form == args.form:
                ms, ph, rt = ms_per_step, phases, routes
            else:
                ms, ph = timed_steps(form, 1, max(1, args.exact_steps) if form == "fp64" else args.steps)
                rt = m.last_stats()
            rec = {"ms_per_step": ms, "ms_median": float(np.median(per_step[form])),
                   "patches_per_s": patches / (ms / 1e3),
                   "phases_ms": [float(x) for x in ph], "routes": rt}
            if form != "fp64":
                sf = spr.getSPRSizeField(eps, 0.05, form=form)
                rel = (sf.data[:, 0] / a64 - 1).abs()
                m.destroyField(sf)
                rf = spr.recoverField(eps, form=form)
                rr = (rf.data - d64).abs()[:, keep] / scale[keep]
                m.destroyField(rf)
                rec.update(size_max_rel=float(rel.max()), size_rms_rel=float(rel.pow(2).mean().sqrt()),
                           recovered_max_rel=float(rr.max()),
                           recovered_rms_rel=float(rr.pow(2).mean().sqrt()))
            forms[form] = rec
            log("form %s: %s" % (form, rec))
        for form in ("fp64f", "fp32r"):
            forms[form]["speedup_vs_fp64_exact"] = forms["fp64"]["ms_per_step"] / forms[form]["ms_per_step"]
        out["forms"] = forms
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
