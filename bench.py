"""Benchmark of the trace-and-field step on one MI355X.

One step = one full sbr_trace of the synthetic urban canyon (482 triangles, icosahedron level 9 =
2,621,442 launches, 6 reflections + 1 transmission, 100x100 receivers): ray generation, the
interaction trees, capture + image correction + duplicate removal, per-path fields and per-receiver
sums, with the scene already resident on the device.  sbr_trace returns after its stream has
drained, so the wall clock around K calls brackets exactly K complete steps.

Metric: Mray-bounces/s = interaction-tree nodes (RayLaunch calls) per second / 1e6.

Prints one JSON line.  --gpus N with N > 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Algorithmic bytes of the closest-hit kernel (k_trace_level): per bounce the parent record is read
# and the new record written (128 B each) plus its queue entry and child mask (8 B); per BVH node
# entered one 64 B node; per ray-triangle test one 96 B triangle record.
B_PER_BOUNCE = 128 + 128 + 8
B_PER_NODE = 64
B_PER_TRI = 96


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(cfg, trace_kw, stride):
    """The FP64 oracle (the reference's algorithm, OpenMP over launches) on every `stride`-th
    launch of the same workload; reports its Mray-bounces/s."""
    from oracle import oracle as O
    sc = O.Scene.from_soup(*cfg["soup"])
    p = O.make_params(**dict(trace_kw, launchStride=stride, launchOffset=0))
    cores = host_cores()
    t0 = time.perf_counter()
    r = O.trace(sc, p, cfg["rx"], cfg["materials"], nthreads=cores)
    dt = time.perf_counter() - t0
    return dict(value=r.bounces / dt / 1e6, unit="Mray-bounces/s", cores=cores, kind="port",
                sample=f"every {stride}th launch ({r.bounces} bounces, {len(cfg['rx'])} receivers), "
                       f"{dt:.1f} s wall, reference traversal + per-receiver tree capture",
                seconds=dt, bounces=int(r.bounces))


def calib_copy(nbytes):
    """One hipMemcpy device-to-device of nbytes (a blit kernel of known traffic: nbytes read and
    nbytes written), for scaling FETCH_SIZE / WRITE_SIZE counter readings."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] calibration copy of {nbytes} bytes done", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tess", type=int, default=9)
    ap.add_argument("--model", type=int, default=1, help="reception model: 1 = alpha*L/sqrt(3), 0 = reference")
    ap.add_argument("--cpu-stride", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--also-ref-model", action="store_true", help="also time the reference reception model")
    ap.add_argument("--calib", action="store_true",
                    help="run one plain 256 MiB device-to-device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design note, not a run",
              file=sys.stderr)
        sys.exit(2)

    from sbrsolver_amd import Context, make_params, scenes
    cfg = scenes.config2(tess=args.tess)
    trace_kw = dict(cfg["trace"], receptionModel=args.model)
    ctx = Context(0)
    ctx.scene_from_soup(*cfg["soup"])
    ctx.materials(cfg["materials"])
    print(f"[bench] scene {cfg['name']} uploaded; warmup {args.warmup}", file=sys.stderr, flush=True)

    # Traversal counts for the algorithmic bytes (one extra, untimed step).
    cnt = ctx.trace(make_params(**trace_kw, countTraversal=1), cfg["rx"]).stats
    for _ in range(args.warmup):
        ctx.trace(make_params(**trace_kw), cfg["rx"])
    params = make_params(**trace_kw)
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats.append(ctx.trace(params, cfg["rx"]).stats)
    dt = time.perf_counter() - t0
    print(f"[bench] {args.steps} steps in {dt:.3f} s", file=sys.stderr, flush=True)

    bounces = stats[-1]["bounces"]
    assert all(s["bounces"] == bounces for s in stats)
    value = bounces * args.steps / dt / 1e6
    ms_step = dt / args.steps * 1e3

    # Dominant kernel and its roofline (kernel times from HIP events on the library's stream).
    k_trace_ms = np.mean([s["ms_k_trace"] for s in stats])
    k_trace_n = stats[-1]["n_k_trace"]
    k_cap_ms = np.mean([s["ms_k_capture"] for s in stats])
    trace_bytes = bounces * B_PER_BOUNCE + cnt["node_visits"] * B_PER_NODE + cnt["tri_tests"] * B_PER_TRI
    stage = {k: float(np.mean([s[k] for s in stats])) for k in ("ms_raygen", "ms_trace", "ms_capture", "ms_field",
                                                                  "ms_total")}
    if k_trace_ms >= k_cap_ms:
        per_launch_bytes = trace_bytes / k_trace_n
        avg_ms = k_trace_ms / k_trace_n
        dom = "k_trace_level"
    else:
        per_launch_bytes = bounces * 128.0   # one tree-node record per node read by the capture kernel
        avg_ms = k_cap_ms
        dom = "k_capture"
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5), traffic=None, kernel=dom,
                    avg_launch_ms=round(avg_ms, 4), algorithmic_bytes_per_launch=int(per_launch_bytes),
                    node_visits=int(cnt["node_visits"]), tri_tests=int(cnt["tri_tests"]),
                    k_trace_ms_per_step=round(float(k_trace_ms), 3), k_trace_launches=int(k_trace_n),
                    k_capture_ms_per_step=round(float(k_cap_ms), 3))

    out = {
        "metric": "Mray-bounces/s",
        "value": round(value, 3),
        "unit": "Mray-bounces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded urban canyon, SplitMix64)",
        "config": {"workload": cfg["name"], "launches": int(stats[-1]["launches"]), "bounces_per_step": int(bounces),
                   "receivers": int(len(cfg["rx"])), "reception_model": "alpha*L/sqrt(3)" if args.model else "reference",
                   "paths_kept": int(stats[-1]["winners"]), "capture_records": int(stats[-1]["capture_records"]),
                   "stage_ms": {k: round(v, 3) for k, v in stage.items()}},
        "roofline": roofline,
    }
    if args.also_ref_model:
        kw0 = dict(cfg["trace"], receptionModel=0)
        ctx.trace(make_params(**kw0), cfg["rx"])
        t1 = time.perf_counter()
        s0 = ctx.trace(make_params(**kw0), cfg["rx"]).stats
        d0 = time.perf_counter() - t1
        out["reference_reception_model"] = dict(value=round(s0["bounces"] / d0 / 1e6, 3), ms_per_step=round(d0 * 1e3, 3),
                                                paths_kept=int(s0["winners"]),
                                                capture_records=int(s0["capture_records"]))
    if args.calib:
        calib_copy(256 << 20)
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(cfg, trace_kw, args.cpu_stride)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
