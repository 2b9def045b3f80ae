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
# Algorithmic bytes of one bounce (one closest-hit query, the work of k_trace):
#   B = 32 (ray in) + 32 (segment out) + 32 * nodes + 48 * triangles,
# with nodes / triangles per bounce those of a canonical closest-first, t-clipped BVH2 traversal
# of the same tree, counted by the oracle over the whole workload and committed as a fixture
# (tests/golden/traversal_counts.json, written by tests/golden/make_traversal_counts.py).
B_RAY_IN, B_SEG_OUT, B_PER_NODE, B_PER_TRI = 32, 32, 32, 48
# VALU model of the same traversal: ~22 FP64 operations per box test, 2 box tests per node
# entered, ~40 per ray-triangle test.
OPS_PER_NODE, OPS_PER_TRI = 44, 40
COUNTS = os.path.join(ROOT, "tests", "golden", "traversal_counts.json")
# HBM traffic of one k_trace launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by the
# readings for a copy of known size (line "traffic_bytes_per_launch: N" in this summary).
TRAFFIC = os.path.join(ROOT, "profiles", "k_trace_traffic_urban_L9_R6T1.md")
# VALU lane activity and wait fraction of the same kernel (rocprofv3 --pmc SQ_* counters).
VALU = os.path.join(ROOT, "profiles", "k_trace_valu_cache_urban_L9_R6T1.md")
# HBM traffic of the bounce kernel on the dense indoor workload (line "traffic_bytes_per_bounce: N").
TRAFFIC_INDOOR = os.path.join(ROOT, "profiles", "k_trace_traffic_indoor_L10.md")


def host_cores():
    """Host cores this process may run on (its affinity mask; all of them are used)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
                       f"{dt:.1f} s wall, reference traversal + per-receiver tree capture, OpenMP over launches "
                       f"on {cores} threads (nproc {os.cpu_count()}, {cpu_model()})",
                seconds=dt, bounces=int(r.bounces))


def cpu_single_thread_box_room():
    """The oracle on one thread, structure-faithful (per-receiver tree clones as RTSolver.cpp:82-121
    does), on the box room in full (12 triangles, level 5 = 10,242 launches, R 4, 100 receivers)."""
    from oracle import oracle as O
    from sbrsolver_amd import scenes
    cfg = scenes.config1()
    sc = O.Scene.from_soup(*cfg["soup"])
    t0 = time.perf_counter()
    r = O.trace(sc, O.make_params(**cfg["trace"]), cfg["rx"], cfg["materials"], nthreads=1)
    dt = time.perf_counter() - t0
    return dict(value=round(r.bounces / dt / 1e6, 5), unit="Mray-bounces/s", cores=1, kind="port",
                workload=cfg["name"], bounces=int(r.bounces), seconds=round(dt, 3), cpu=cpu_model())


def kernel_counters(traffic, avg_ms, valu_tflops):
    """What bounds k_trace, from the committed rocprofv3 counter summaries (profiles/): the measured
    HBM traffic as a fraction of peak bandwidth, the FP64 VALU model as a fraction of peak, the
    lanes active per VALU instruction, the fractions of wave cycles spent waiting on results and
    waiting to issue, and the share of the SIMDs' VALU issue slots the kernel's instructions take.
    The algorithmic bytes of the roofline line stay in LDS and the caches (the counters see a fifth
    of them in HBM), so neither HBM nor the FP64 peak binds: the kernel is bound by the latency of
    its dependent traversal steps and by instruction issue (about half of the VALU slots, a quarter
    of the wave cycles waiting for issue) at its occupancy."""
    out = {}
    if traffic:
        out["hbm_measured_GBs"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
        out["hbm_measured_frac"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    out["valu_fp64_model_frac"] = round(valu_tflops / FP64_PEAK_TFLOPS, 4)
    if os.path.exists(VALU):
        import re
        txt = open(VALU).read()
        def num(name):
            m = re.search(r"\| " + name + r" \| ([0-9,]+) \|", txt)
            return float(m.group(1).replace(",", "")) if m else None
        tc, iv, wait, wc, wi = (num("SQ_THREAD_CYCLES_VALU"), num("SQ_INSTS_VALU"), num("SQ_WAIT_ANY"),
                                num("SQ_WAVE_CYCLES"), num("SQ_WAIT_INST_ANY"))
        if tc and iv:
            out["valu_active_lanes_frac"] = round(tc / iv / 64.0, 4)
        if wait and wc:
            out["wave_wait_frac"] = round(wait / wc, 4)
        if wi and wc:
            out["wave_wait_issue_frac"] = round(wi / wc, 4)
        m = re.search(r"valu_issue_frac:\s*([0-9.]+)", txt)
        if m:
            out["valu_issue_frac"] = float(m.group(1))
    fr = [v for k, v in out.items() if k in ("hbm_measured_frac", "valu_fp64_model_frac")]
    nan = float("nan")
    out["binding"] = ("latency and issue (waves wait {:.0%} of their cycles on results and {:.0%} to issue; VALU "
                      "slots {:.0%} busy; HBM {:.0%} and FP64 VALU {:.0%} of peak)".format(
                          out.get("wave_wait_frac", nan), out.get("wave_wait_issue_frac", nan),
                          out.get("valu_issue_frac", nan), out.get("hbm_measured_frac", nan),
                          out["valu_fp64_model_frac"])
                      if fr and max(fr) < 0.5 else ("hbm" if out.get("hbm_measured_frac", 0) >= out["valu_fp64_model_frac"]
                                                    else "valu"))
    return out


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


def batched_run(ctx, make_params, scenes):
    """Batched multi-transmitter workload (256 transmitters on the urban scene, levels 6-9, 4-6
    reflections, 10k shared receivers): one sbr_trace_batch call over all transmitters (the library
    runs it in launch windows of its tree-node budget), against one sbr_trace per transmitter on the
    same inputs."""
    cfg = scenes.config4()
    plist = [make_params(**t) for t in cfg["txs"]]
    windows = []

    def run_batched():
        st = ctx.trace_batch(plist, cfg["rx"]).stats
        windows[:] = [st["windows"]]
        return st["bounces"]

    def run_single():
        b = 0
        for p in plist:
            b += ctx.trace(p, cfg["rx"]).stats["bounces"]
        return b

    run_batched()
    t0 = time.perf_counter()
    bb = run_batched()
    tb = time.perf_counter() - t0
    run_single()
    t0 = time.perf_counter()
    bs = run_single()
    ts = time.perf_counter() - t0
    assert bb == bs
    return dict(workload=cfg["name"], transmitters=len(plist), batches=1, windows=int(windows[0]),
                launches=int(sum(10 * 4 ** p.tess + 2 for p in plist)), bounces=int(bb),
                batched_mrays=round(bb / tb / 1e6, 3), batched_s=round(tb, 4), single_tx_mrays=round(bs / ts / 1e6, 3),
                single_tx_s=round(ts, 4), speedup=round(ts / tb, 3))


def indoor_run(Context, make_params, scenes, tess=10):
    """Dense indoor workload (about 207k triangles, icosahedron level 10, reflections and
    transmissions up to 8 each and 8 in total, 20k receivers) on a context of its own: one warm-up
    trace, one timed, then one counting pass (the kernel's own traversal counts).  Beside them: the
    reference traversal's canonical counts (committed fixture, every 64th launch) and the bounce
    kernel's measured HBM traffic (committed rocprofv3 summary) as a fraction of the peak."""
    cfg = scenes.config3(tess=tess)
    with Context(0) as ictx:
        ictx.scene_from_soup(*cfg["soup"])
        ictx.materials(cfg["materials"])
        p = make_params(**cfg["trace"])
        ictx.trace(p, cfg["rx"])
        t0 = time.perf_counter()
        s = ictx.trace(p, cfg["rx"]).stats
        dt = time.perf_counter() - t0
        c = ictx.trace(make_params(**dict(cfg["trace"], countTraversal=1)), cfg["rx"]).stats
    b = s["bounces"]
    out = dict(workload=cfg["name"], triangles=int(len(cfg["soup"][2])), launches=int(s["launches"]),
               bounces=int(b), windows=int(s["windows"]), value=round(b / dt / 1e6, 3),
               unit="Mray-bounces/s", s_per_trace=round(dt, 3), ms_trace=round(s["ms_trace"], 1),
               ms_k_trace=round(s["ms_k_trace"], 1), k_trace_launches=int(s["n_k_trace"]),
               ms_capture=round(s["ms_capture"], 1), paths_kept=int(s["winners"]),
               kernel_nodes_per_bounce=round(c["node_visits"] / b, 3),
               kernel_exact_tri_tests_per_bounce=round(c["tri_tests"] / b, 3))
    with open(COUNTS) as f:
        canon = json.load(f).get(cfg["name"])
    if canon:
        out["reference_nodes_per_bounce"] = round(canon["nodes_per_bounce"], 3)
        out["reference_tris_per_bounce"] = round(canon["tris_per_bounce"], 3)
        out["reference_counts_sample"] = f"every {canon['launch_stride']}th launch ({canon['bounces']} bounces)"
    if os.path.exists(TRAFFIC_INDOOR):
        import re
        m = re.search(r"traffic_bytes_per_bounce:\s*([0-9.eE+]+)", open(TRAFFIC_INDOOR).read())
        if m:
            tb = float(m.group(1))
            gbs = tb * b / (s["ms_k_trace"] * 1e-3) / 1e9
            out["hbm"] = dict(bytes_per_bounce=round(tb, 1), achieved_GBs=round(gbs, 1), peak_GBs=HBM_PEAK_GBS,
                              frac=round(gbs / HBM_PEAK_GBS, 4),
                              note="bounce kernel's HBM traffic per bounce (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                   "scaled by a copy of known size) over its event-timed duration")
    return out


def serve_run(Context, make_params, scenes, Server, n=96, slots=3, load=0.7):
    """Serving stream on the seeded city (about 1M triangles, 2000 buildings): the scene is built and
    made resident once (its load time is reported), then n coverage queries (a transmitter in a
    street at 3-30 m, levels 6-10, 4 reflections, 1..50k receivers over the central 1.35 x 1.35 km).
    Phase 1 submits all at once to measure capacity (req/s when saturated); phase 2 submits the same
    queries with Poisson arrivals at `load` x capacity and reports the latency (submit -> fields on
    the host) of that phase."""
    import random
    t0 = time.perf_counter()
    name, soup, boxes, half, mats = scenes.config5_scene()
    t_gen = time.perf_counter() - t0
    ctx = Context(0)
    t0 = time.perf_counter()
    ms_dev = ctx.scene_from_soup_device(*soup)   # triangles, coplanar ids and a linear BVH built on the device
    ctx.materials(mats)
    t_load = time.perf_counter() - t0
    print(f"[bench] serving: {name} generated in {t_gen:.1f} s, loaded in {t_load:.2f} s", file=sys.stderr, flush=True)
    qs = scenes.coverage_queries(n, boxes, half=0.5 * half)
    params = [make_params(**kw) for kw, _ in qs]

    def finish(srv, ids):
        infos = []
        for q in ids:
            rc, info = srv.poll(q, wait_ms=-1)
            if rc != 0:
                raise RuntimeError(f"query failed: {srv.error(q)}")
            infos.append(info)
            srv.release(q)
        return infos

    with Server(ctx, slots=slots) as srv:
        finish(srv, [srv.submit(p, rx) for p, (_, rx) in zip(params, qs)])      # warm engines and ray caches
        t0 = time.perf_counter()
        infos = finish(srv, [srv.submit(p, rx) for p, (_, rx) in zip(params, qs)])
        t_sat = time.perf_counter() - t0
        capacity = n / t_sat
        rate = load * capacity
        rnd = random.Random(20261203 + 5)
        ids = []
        t0 = time.perf_counter()
        t_next = t0
        for p, (_, rx) in zip(params, qs):
            t_next += rnd.expovariate(rate)
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(p, rx))
        infos = finish(srv, ids)
        t_poisson = time.perf_counter() - t0
        st = srv.stats()
    ctx.close()
    lat = np.sort([i["ms_total"] for i in infos])
    bounces = sum(i["bounces"] for i in infos)
    nearest = lambda q: float(lat[min(len(lat), max(1, int(np.ceil(q * len(lat))))) - 1])
    return dict(workload=f"{name}_coverage_queries_L6-10_R4_rx1-50k", queries=n, slots=slots,
                triangles=int(len(soup[2])), scene_generate_s=round(t_gen, 2), scene_load_s=round(t_load, 3),
                scene_device_build_ms=round(ms_dev[0], 1),
                scene_load_note="sbr_scene_from_soup_device: vertex dedup, coplanar ids and a linear BVH on the device, then the upload",
                capacity_req_s=round(capacity, 2), capacity_mray_bounces_s=round(bounces / t_sat / 1e6, 2),
                offered_req_s=round(rate, 2), req_per_s=round(n / t_poisson, 2),
                mray_bounces_per_s=round(bounces / t_poisson / 1e6, 2), p50_ms=round(nearest(0.5), 3),
                p99_ms=round(nearest(0.99), 3), mean_ms=round(float(lat.mean()), 3),
                max_ms=round(float(lat[-1]), 3), mean_queued_ms=round(float(np.mean([i["ms_queued"] for i in infos])), 3),
                peak_inflight_bytes=int(st["peak_inflight_bytes"]), budget_bytes=int(st["budget_bytes"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tess", type=int, default=9)
    ap.add_argument("--model", type=int, default=1, help="reception model: 1 = alpha*L/sqrt(3), 0 = reference")
    ap.add_argument("--cpu-stride", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ref-model", action="store_true", help="skip the reference reception model line")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-transmitter line")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop line")
    ap.add_argument("--no-indoor", action="store_true", help="skip the dense indoor line")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched multi-transmitter workload and print its line (profiling aid)")

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
    if args.batch_only:
        print(json.dumps(dict(batched=batched_run(ctx, make_params, scenes))), flush=True)
        ctx.close()
        return

    # One cold step: the first trace of the context generates the level's launch directions (later
    # steps reuse them) and sizes its buffers.
    t0 = time.perf_counter()
    cold = ctx.trace(make_params(**trace_kw), cfg["rx"]).stats
    cold_step = dict(ms_wall=round((time.perf_counter() - t0) * 1e3, 3), ms_raygen=round(cold["ms_raygen"], 3),
                     ms_total=round(cold["ms_total"], 3),
                     note="first trace of a fresh context: ray generation and buffer allocation included")

    # The kernel's own traversal counts (one extra, untimed step; reported beside the canonical ones).
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

    # Dominant kernel (k_trace: one launch per tree depth) and its roofline.  Kernel time is the
    # HIP-event time of the launches on the library's stream, averaged per launch.
    k_trace_ms = float(np.mean([s["ms_k_trace"] for s in stats]))
    k_trace_n = int(stats[-1]["n_k_trace"])
    k_cap_ms = float(np.mean([s["ms_k_capture"] for s in stats]))
    stage = {k: float(np.mean([s[k] for s in stats])) for k in ("ms_raygen", "ms_trace", "ms_capture", "ms_field",
                                                                  "ms_total")}
    with open(COUNTS) as f:
        canon = json.load(f)[cfg["name"]]
    assert canon["bounces"] == bounces, "committed traversal counts are for another workload"
    b_bounce = B_RAY_IN + B_SEG_OUT + B_PER_NODE * canon["nodes_per_bounce"] + B_PER_TRI * canon["tris_per_bounce"]
    per_launch_bytes = bounces * b_bounce / k_trace_n
    avg_ms = k_trace_ms / k_trace_n
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC):
        import re
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(TRAFFIC).read())
        traffic = float(m.group(1)) if m else None
    ops = bounces * (OPS_PER_NODE * canon["nodes_per_bounce"] + OPS_PER_TRI * canon["tris_per_bounce"])
    valu_tflops = ops / (k_trace_ms * 1e-3) / 1e12
    counters = kernel_counters(traffic, avg_ms, valu_tflops)
    roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic, kernel="k_trace",
                    binding=counters.pop("binding"), counters=counters,
                    avg_launch_ms=round(avg_ms, 4), launches_per_step=k_trace_n,
                    bytes_per_bounce=round(b_bounce, 1), algorithmic_bytes_per_launch=int(per_launch_bytes),
                    canonical_nodes_per_bounce=round(canon["nodes_per_bounce"], 4),
                    canonical_tris_per_bounce=round(canon["tris_per_bounce"], 4),
                    kernel_nodes_per_bounce=round(cnt["node_visits"] / bounces, 4),
                    kernel_exact_tri_tests_per_bounce=round(cnt["tri_tests"] / bounces, 4),
                    capture_sphere_tests_per_bounce=round(cnt["capture_tests"] / bounces, 4),
                    capture_nodes_per_bounce=round(cnt["capture_nodes"] / bounces, 4),
                    fp64_box_tests_per_bounce=round(cnt["box_tests_fp64"] / bounces, 4),
                    wide_redone_per_bounce=round(cnt["wide_redone"] / bounces, 7),
                    valu_model=dict(achieved=round(valu_tflops, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s (FP64)",
                                    frac=round(valu_tflops / FP64_PEAK_TFLOPS, 5)),
                    k_capture_ms_per_step=round(k_cap_ms, 3))

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
        "cold_step": cold_step,
    }
    def extra(name, fn):
        """Secondary lines: a failure is reported in the line instead of failing the headline."""
        print(f"[bench] {name}", file=sys.stderr, flush=True)
        try:
            out[name] = fn()
        except Exception as e:   # noqa: BLE001
            out[name] = {"error": f"{type(e).__name__}: {e}"}

    if not args.no_ref_model:
        def ref_model():
            kw0 = dict(cfg["trace"], receptionModel=0)
            ctx.trace(make_params(**kw0), cfg["rx"])
            t1 = time.perf_counter()
            s0 = ctx.trace(make_params(**kw0), cfg["rx"]).stats
            d0 = time.perf_counter() - t1
            return dict(value=round(s0["bounces"] / d0 / 1e6, 3), unit="Mray-bounces/s", ms_per_step=round(d0 * 1e3, 3),
                        paths_kept=int(s0["winners"]), capture_records=int(s0["capture_records"]))
        extra("reference_reception_model", ref_model)
    if not args.no_batch:
        extra("batched", lambda: batched_run(ctx, make_params, scenes))
    if not args.no_serve:
        from sbrsolver_amd import Server
        extra("serving", lambda: serve_run(Context, make_params, scenes, Server))
    if not args.no_indoor:
        extra("indoor", lambda: indoor_run(Context, make_params, scenes))

    # Batched throughput as a fraction of the same HBM roofline: its bounces priced at the headline
    # workload's algorithmic bytes per bounce (same scene; the per-bounce traversal counts of its
    # levels are not committed, so this is a model figure).  The served city is priced with its own
    # committed canonical counts (the reference traversal over the query stream, sampled).
    if "batched" in out and "batched_mrays" in out["batched"]:
        # the second denominator: the plain single-transmitter kernel on the headline workload
        out["batched"]["speedup_vs_headline_single_tx"] = round(out["batched"]["batched_mrays"] / value, 3)
    for name, key in (("batched", "batched_mrays"),):
        if name in out and key in out[name]:
            gbs = out[name][key] * 1e6 * b_bounce / 1e9
            out[name]["roofline_frac_" + key] = round(gbs / HBM_PEAK_GBS, 5)
    if "serving" in out and "capacity_mray_bounces_s" in out["serving"]:
        with open(COUNTS) as f:
            city = json.load(f).get("city")
        if city:
            bc = B_RAY_IN + B_SEG_OUT + B_PER_NODE * city["nodes_per_bounce"] + B_PER_TRI * city["tris_per_bounce"]
            sv = out["serving"]
            sv["bytes_per_bounce"] = round(bc, 1)
            for key in ("capacity_mray_bounces_s", "mray_bounces_per_s"):
                sv["roofline_frac_" + key] = round(sv[key] * 1e6 * bc / 1e9 / HBM_PEAK_GBS, 5)

    if args.calib:
        calib_copy(256 << 20)
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(cfg, trace_kw, args.cpu_stride)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        print("[bench] cpu baseline, single thread, box room", file=sys.stderr, flush=True)
        out["cpu_baseline_single_thread"] = cpu_single_thread_box_room()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
