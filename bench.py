"""Bounce-loop benchmark: converged bounces/s on one MI355X.

Workload (default): the Sinai-like scene -- 4 spheres + 1 ellipsoid, 2,048
order-4 (5x5 tensor) patches each, 10,240 patches -- and 2^20 rays from the
z=-8 plane in a 10 degree cone, max 100 bounces, FP64 Newton in the
`converged` mode, triangle (stage-1) seeds.  One step = one trace of the whole
ray batch, inputs and outputs resident in HBM.  By default four steps are kept
in flight on four streams (bbrt_trace_device_async), each with its own outputs:
the next step's rays start on the CUs the previous step's last rays leave.  A
pass with one synchronous trace at a time is timed too (`one_in_flight`) and
gives the roofline's kernel launch times.

Prints one JSON line with the throughput, the FP64 roofline of the bounce
kernel (algorithmic Newton flops / summed kernel time, HIP events on the
stream the kernels run on) and the CPU oracle timed on a bounded sample of the
same rays on the host cores.  Two more legs ride along in the same line: the
learned Newton seed (trained on the GPU first, then the same trace) and the
batched bounce map (256 independent scenes in one hipGraph-replayed launch).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Streams that share a hardware queue run their kernels in order.  With the
# HIP default of 4 queues per process, the first streams created after a scene
# handle's own two share queues, which would serialise the in-flight traces;
# 8 queues give the handle's streams and four in-flight streams one each.  Set
# before the HIP runtime starts.
os.environ["GPU_MAX_HW_QUEUES"] = "8"

METRIC = "converged bounces/s at ≤1e-10 hit error; mean Newton iters, learned vs tri seed"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rays", type=int, default=1 << 20)
    ap.add_argument("--max-bounces", type=int, default=100)
    ap.add_argument("--record-k", type=int, default=-1, help="bounces recorded per ray (-1 = max bounces)")
    ap.add_argument("--mode", choices=["ref", "converged"], default="converged")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU oracle time budget (0 = skip)")
    ap.add_argument("--in-flight", type=int, default=4,
                    help="Sinai headline: traces kept in flight on that many streams (bbrt_trace_device_async); "
                         "1 = one synchronous trace at a time")
    ap.add_argument("--in-flight-order", choices=["grouped", "input"], default="grouped",
                    help="queue order of the in-flight passes (Sinai and batched legs): BBRT_FLAG_GROUP_RAYS or input")
    ap.add_argument("--no-cull", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--waves-per-eu", type=int, default=0)
    ap.add_argument("--wavefront", action="store_true", help="per-bounce launches instead of the persistent kernel")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-seed leg")
    ap.add_argument("--no-seed-first", action="store_true",
                    help="learned leg: skip the network-first pass (BBRT_FLAG_SEED_FIRST)")
    ap.add_argument("--train-samples", type=int, default=20_000_000)
    ap.add_argument("--train-epochs", type=int, default=10)
    ap.add_argument("--train-batch", type=int, default=8192)
    ap.add_argument("--train-lr", type=float, default=3e-3)
    ap.add_argument("--newton-batch", type=int, default=0, help="lanes that run Newton together in the persistent kernel (0 = library default)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-in-flight", type=int, default=4,
                    help="batched leg: traces kept in flight (bbrt_trace_device_async); 1 = graph replays only")
    ap.add_argument("--batch-no-tune", action="store_true", help="batched leg: library default launch shape")
    ap.add_argument("--batch-eager", action="store_true",
                    help="batched leg: eager launches instead of hipGraph replays (for counter profiles)")
    ap.add_argument("--batch-no-check", dest="batch_check", action="store_false",
                    help="batched leg: skip the CPU oracle check of sampled scenes")
    ap.add_argument("--batch-only", action="store_true", help="run the batched-scenes leg alone")
    ap.add_argument("--workload", choices=["sinai", "full-form"], default="sinai",
                    help="full-form: the full batched form (512 scenes, 16M rays, 200 bounces, learned seeds) alone")
    ap.add_argument("--no-full-form", action="store_true", help="skip the full batched form leg (512 scenes, 16M rays)")
    ap.add_argument("--full-newton", choices=["fp32", "fp64"], default="fp32",
                    help="full form: Newton's first iterations in FP32 with the decisions and polish in FP64 "
                         "(BBRT_NEWTON_CONVERGED_FP32, the full form's precision split), or FP64 throughout")
    ap.add_argument("--full-no-other-newton", dest="full_other_newton", action="store_false",
                    help="full form: skip the learned-seed pass in the other Newton precision split")
    ap.add_argument("--full-no-triangle", dest="full_triangle", action="store_false",
                    help="full form: skip the triangle-seed run of the same workload")
    ap.add_argument("--full-scenes", type=int, default=512)
    ap.add_argument("--full-rays", type=int, default=16_000_000)
    ap.add_argument("--full-in-flight", type=int, default=2,
                    help="full form: traces kept in flight (bbrt_trace_device_async); 1 = graph replays only")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a device copy of this many MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe_mb > 0:
        copy_probe(args.copy_probe_mb)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the bounce loop here is built and timed on a single MI355X "
              "(rays are independent; a multi-GPU run would shard rays and all-reduce hit counts).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from billardballraytracing3d_amd import scenes as S, tracer as T, _lib as L

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    if args.workload == "full-form":
        r = full_form_leg(args)
        print(json.dumps(r), flush=True)
        return
    if args.batch_only:
        b = batch_leg(args)
        print(json.dumps({"metric": "converged bounces/s", "value": b["value"], "unit": "bounces/s", "n_gpus": 1,
                          "steps": b["steps"], "warmup": args.warmup, "ms_per_step": b["ms_per_step"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic", "config": {"workload": b["workload"]}, "batch": b}), flush=True)
        return
    t0 = time.time()
    sc, o, d, mb = S.sinai_cone(seed=args.seed, n_rays=args.rays)
    mb = args.max_bounces
    K = mb if args.record_k < 0 else args.record_k
    gs = T.GpuScene(sc)
    t_build = time.time() - t0
    n = o.shape[0]
    mode = L.NEWTON_CONVERGED if args.mode == "converged" else L.NEWTON_REF
    flags = L.FLAG_TIME_KERNELS | (L.FLAG_NO_CULL if args.no_cull else 0) | (L.FLAG_WAVEFRONT if args.wavefront else 0)

    to = torch.from_numpy(o).to(dev)
    td = torch.from_numpy(d).to(dev)
    outs = dict(status=torch.empty(n, dtype=torch.int8, device=dev),
                err_kind=torch.empty(n, dtype=torch.int8, device=dev),
                bounces=torch.empty(n, dtype=torch.int32, device=dev),
                err_patch=torch.empty(n, dtype=torch.int32, device=dev),
                final_od=torch.empty((n, 6), dtype=torch.float64, device=dev))
    if K > 0:
        outs.update(hit_patch=torch.empty((n, K), dtype=torch.int32, device=dev),
                    hit_xyz=torch.empty((n, K, 3), dtype=torch.float64, device=dev),
                    newton_iters=torch.empty((n, K), dtype=torch.uint8, device=dev))
    ptrs = {k: v.data_ptr() for k, v in outs.items()}
    stream = torch.cuda.current_stream(dev)

    def step():
        return gs.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, mode=mode, record_k=K, flags=flags,
                               stream=stream.cuda_stream, block=args.block, waves_per_eu=args.waves_per_eu,
                               newton_batch=args.newton_batch)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctr_tot = np.zeros(L.NCOUNTERS, np.int64)
    kms_tot = np.zeros(3)
    ts = time.perf_counter()
    for _ in range(args.steps):
        c, k = step()
        ctr_tot += c
        kms_tot += k
    torch.cuda.synchronize()
    te = time.perf_counter()
    ms_per_step = (te - ts) * 1e3 / args.steps
    bounces = int(ctr_tot[L.CTR["bounces"]])
    value = bounces / (te - ts)
    launches = kms_tot[1]
    kernel_ms = kms_tot[0]
    flops = float(ctr_tot[L.CTR["newton_flops"]]) + 27.0 * bounces
    achieved_tflops = flops / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0
    # The pass above runs one synchronous trace at a time: its HIP-event launch times give the roofline.
    # The headline then keeps in_flight traces in flight on as many streams (bbrt_trace_device_async), so a
    # step's first rays start on the CUs the previous step's last rays have left; every step is the whole
    # trace of all 2^20 rays, with its own outputs and device counters.
    one_in_flight = {"value": value, "ms_per_step": ms_per_step, "steps": args.steps,
                     "note": "one synchronous trace at a time; the roofline's launch times come from this pass"}
    in_flight = 1 if args.wavefront else max(1, args.in_flight)
    in_flight_check = None
    if in_flight > 1:
        dt, cd, in_flight_check = run_in_flight(
            gs, to, td, n, mb, outs, args.steps, args.warmup, in_flight, ctr_tot // max(1, args.steps),
            dict(mode=mode, record_k=K, flags=(flags & ~L.FLAG_TIME_KERNELS) | group_flag(args), block=args.block,
                 waves_per_eu=args.waves_per_eu, newton_batch=args.newton_batch))
        bounces = int(cd[:, L.CTR["bounces"]].sum())
        value = bounces / dt
        ms_per_step = dt * 1e3 / args.steps
    newton_calls = int(ctr_tot[L.CTR["newton_calls"]])
    newton_iters = int(ctr_tot[L.CTR["newton_iters"]])
    per_step = {k: int(ctr_tot[v] // args.steps) for k, v in L.CTR.items()}

    res = {
        "metric": METRIC,
        "value": value,
        "unit": "bounces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "sinai-5-obstacles-p4-10240-patches", "rays": n, "patches": sc.n_patches,
                   "traces_in_flight": in_flight,
                   "queue_order": (args.in_flight_order if in_flight > 1 else "input"),
                   "max_bounces": mb, "record_k": K, "newton_mode": args.mode, "seed_mode": "triangle",
                   "culling": not args.no_cull, "scene_seed": args.seed,
                   "schedule": "wavefront" if args.wavefront else "persistent", "block": args.block or 64,
                   "waves_per_eu": args.waves_per_eu or 2, "newton_batch": args.newton_batch or 64},
        "roofline": {"bound": "mfma", "compute": "fp64 (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
                     "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_bounce" if args.wavefront else "k_persist", "launches_per_step": launches / args.steps,
                     "avg_launch_ms": kernel_ms / launches if launches else None,
                     "flops_per_launch": flops / launches if launches else None,
                     "pass": "one synchronous trace at a time (each launch's own duration, HIP events)",
                     "per_headline_step": {
                         "achieved": (flops / launches) / (ms_per_step * 1e-3) / 1e12 if launches else None,
                         "frac": (flops / launches) / (ms_per_step * 1e-3) / 1e12 / FP64_PEAK_TFLOPS if launches else None,
                         "note": "the same flops per trace over the headline's ms_per_step (traces overlap in flight)"}},
        "one_in_flight": one_in_flight,
        "in_flight_check": in_flight_check,
        "mean_newton_iters": {"triangle": newton_iters / newton_calls if newton_calls else None, "learned": None},
        "bounces_per_step": bounces // args.steps,
        "rays_by_status": status_counts(outs["status"]),
        "errors_by_kind": error_counts(outs["err_kind"]),
        "counters_per_step": per_step,
        "scene_build_s": t_build,
    }

    if not args.no_learned:
        if mode != L.NEWTON_CONVERGED or args.wavefront:
            res["learned"] = {"skipped": "the learned seed runs with converged Newton in the persistent schedule"}
        else:
            res["learned"] = learned_leg(args, gs, to, td, n, mb, K, outs, ptrs, stream, flags,
                                         one_in_flight["value"], res)
            res["mean_newton_iters"]["learned"] = res["learned"]["mean_newton_iters"]

    traffic = committed_profile_value(args, n, mb, K, "traffic_bytes_per_launch")
    avg_ms = kernel_ms / launches if launches else None
    if traffic is not None:
        res["roofline"]["traffic"] = traffic
        res["roofline"]["traffic_source"] = "profiles/k_persist_sinai_1M.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-scaled)"
        if avg_ms:
            res["roofline"]["hbm"] = {"achieved": traffic / (avg_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    lane_flops = committed_profile_value(args, n, mb, K, "fp64_valu_lane_flops_per_launch")
    if lane_flops is not None and avg_ms:
        res["roofline"]["measured_fp64_valu"] = {
            "lane_flops_per_launch": lane_flops, "achieved": lane_flops / (avg_ms * 1e-3) / 1e12,
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": lane_flops / (avg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "source": "profiles/k_persist_sinai_1M.md (rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes; issued lane "
                      "slots, masked lanes included)"}
    if args.cpu_seconds > 0:
        res["cpu_baseline"] = cpu_baseline(sc, o, d, mb, args.cpu_seconds, gs, K)

    # the batched legs need the HBM the Sinai scene and its buffers hold
    del to, td, outs, ptrs, gs
    torch.cuda.empty_cache()
    if not args.no_batch:
        _progress("batched bounce map leg")
        res["batch"] = batch_leg(args)
        torch.cuda.empty_cache()
    if not args.no_full_form:
        _progress("full batched form leg")
        res["full_form"] = full_form_leg(args)
    print(json.dumps(res), flush=True)


def batch_leg(args):
    """The batched bounce map: args.batch_scenes independent scenes (1-8 obstacles, patches log-uniform in
    [80, 5120] per obstacle, order 2/4, 1k-64k rays each, 30 degree cone, 50 bounces) in one compact scene
    handle built on the device, launch shape autotuned, traced by one hipGraph-replayed persistent launch."""
    import numpy as np
    import torch
    from billardballraytracing3d_amd import scenes as S, tracer as T, _lib as L

    dev = torch.device("cuda", 0)
    t0 = time.time()
    scenes, o, d, rb = S.batch_config(args.seed, n_scenes=args.batch_scenes)
    t1 = time.time()
    gb = T.GpuScene.batch(scenes, compact=True, device_build=True)
    t2 = time.time()
    n_patches = gb.n_patches
    sampled = {int(i): scenes[int(i)] for i in np.linspace(0, len(scenes) - 1, 6).astype(int)}
    del scenes
    n = int(rb[-1])
    mb, K = 50, 20
    to = torch.from_numpy(o).to(dev)
    td = torch.from_numpy(d).to(dev)
    outs = dict(status=torch.empty(n, dtype=torch.int8, device=dev),
                err_kind=torch.empty(n, dtype=torch.int8, device=dev),
                bounces=torch.empty(n, dtype=torch.int32, device=dev),
                hit_patch=torch.empty((n, K), dtype=torch.int32, device=dev),
                hit_xyz=torch.empty((n, K, 3), dtype=torch.float64, device=dev))
    ptrs = {k: v.data_ptr() for k, v in outs.items()}
    st = torch.cuda.current_stream(dev).cuda_stream
    ta = time.time()
    if args.batch_no_tune:
        best, rows = {}, []
    else:
        best, rows = gb.autotune(to.data_ptr(), td.data_ptr(), n, mb, ptrs, record_k=K, ray_begin=rb, stream=st)
    t_tune = time.time() - ta

    graph_flag = 0 if args.batch_eager else L.FLAG_GRAPH

    def step():
        return gb.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, record_k=K, ray_begin=rb, stream=st,
                               flags=graph_flag | (L.FLAG_TIME_KERNELS if args.batch_eager else 0), **best)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    ctr = np.zeros(L.NCOUNTERS, np.int64)
    kms = np.zeros(3)
    ts = time.perf_counter()
    for _ in range(args.batch_steps):
        c, k = step()
        ctr += c
        kms += k
    torch.cuda.synchronize()
    te = time.perf_counter()
    bounces = int(ctr[L.CTR["bounces"]])
    flops = float(ctr[L.CTR["newton_flops"]]) + 27.0 * bounces
    graph_ms = kms[0] / args.batch_steps
    one_in_flight = {"value": bounces / (te - ts), "ms_per_step": (te - ts) * 1e3 / args.batch_steps,
                     "note": "one eager trace at a time" if args.batch_eager else "one hipGraph-replayed trace at a time"}
    check = oracle_check(sampled, o, d, rb, outs, mb, K, "batch") if args.batch_check else None
    in_flight_check = None
    if args.batch_in_flight > 1:
        dt, cd, in_flight_check = run_in_flight(
            gb, to, td, n, mb, outs, args.batch_steps, max(1, args.warmup), args.batch_in_flight,
            ctr // args.batch_steps, dict(record_k=K, ray_begin=rb, flags=group_flag(args), **best))
        ts, te = 0.0, dt
    return {
        "workload": f"batched-bounce-map-{args.batch_scenes}-scenes", "value": bounces / (te - ts),
        "unit": "bounces/s", "ms_per_step": (te - ts) * 1e3 / args.batch_steps, "steps": args.batch_steps,
        "traces_in_flight": max(1, args.batch_in_flight), "one_in_flight": one_in_flight,
        "in_flight_check": in_flight_check,
        "scenes": args.batch_scenes, "patches": n_patches, "rays": n, "max_bounces": mb, "record_k": K,
        "newton_mode": "converged", "scene": "compact (coarse cells evaluated on the fly), device build",
        "bounces_per_step": bounces // args.batch_steps, "graph_ms_per_step": graph_ms,
        "fp64_tflops": flops / (graph_ms * args.batch_steps * 1e-3) / 1e12 if graph_ms > 0 else None,
        "autotuned": best, "autotune_candidates": rows, "autotune_s": t_tune, "oracle_check": check,
        "scene_generation_s": t1 - t0, "scene_build_s": t2 - t1,
        "counters_per_step": {k: int(ctr[v] // args.batch_steps) for k, v in L.CTR.items()},
    }


ERROR_KINDS = ("none", "newton_singular", "newton_noconv", "newton_domain", "backface", "zero_dir", "stack")


def group_flag(args):
    """BBRT_FLAG_GROUP_RAYS for the in-flight passes unless --in-flight-order input: grouped rays make each
    wave's rays coherent, which pays once traces overlap (measured +3.6 % on the Sinai workload)."""
    from billardballraytracing3d_amd import _lib as L
    return L.FLAG_GROUP_RAYS if args.in_flight_order == "grouped" else 0


def run_in_flight(g, to, td, n, mb, outs, steps, warmup, in_flight, ref_counters, trace_kw):
    """Times `steps` traces kept `in_flight` deep on as many streams (bbrt_trace_device_async).  Each stream
    writes its own outputs (the first set is `outs`), each trace its own device counters.  Returns (seconds,
    counters [steps][NCOUNTERS], check); exits if a trace's counters differ from ref_counters (a synchronous
    trace's) or an output set differs from `outs`."""
    import torch
    from billardballraytracing3d_amd import _lib as L

    streams = [torch.cuda.Stream(to.device) for _ in range(in_flight)]
    extra = [{k: torch.empty_like(v) for k, v in outs.items()} for _ in range(in_flight - 1)]
    sets = [{k: v.data_ptr() for k, v in outs.items()}] + [{k: v.data_ptr() for k, v in e.items()} for e in extra]
    ctr = torch.zeros((max(1, steps), L.NCOUNTERS), dtype=torch.int64, device=to.device)

    def astep(i):
        g.trace_device_async(to.data_ptr(), td.data_ptr(), n, mb, sets[i % in_flight],
                             counters_ptr=ctr[i % ctr.shape[0]].data_ptr(), stream=streams[i % in_flight].cuda_stream,
                             **trace_kw)

    for i in range(max(warmup, in_flight)):
        astep(i)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for i in range(steps):
        astep(i)
    torch.cuda.synchronize()
    dt = time.perf_counter() - ts
    cd = ctr.cpu().numpy()[:steps]
    # launches and the network's wave-level evaluation counts depend on scheduling, not on the rays
    sched = ("launches", "mlp_evals", "mlp_lanes")
    differ = [name for name, q in L.CTR.items() if name not in sched and bool((cd[:, q] != ref_counters[q]).any())]
    def same_outputs(a, b):
        # per-bounce record slots past a ray's bounce count are left unwritten by the device entry points
        ok = True
        for k in a:
            if k in ("hit_patch", "hit_xyz", "newton_iters") and "bounces" in a:
                K = a[k].shape[1]
                slot = torch.arange(K, device=a[k].device)[None, :]
                valid = slot < torch.clamp(b["bounces"], max=K)[:, None].to(slot.dtype)
                if a[k].dim() == 3:
                    valid = valid[:, :, None]
                ok = ok and bool(torch.all((a[k] == b[k]) | ~valid).item())
            else:
                ok = ok and torch.equal(a[k], b[k])
        return ok

    check = {"traces_in_flight": in_flight, "counters_equal_synchronous": not differ,
             "outputs_equal_across_streams": all(same_outputs(e, outs) for e in extra)}
    if differ:
        check["counters_differing"] = differ
    if not (check["counters_equal_synchronous"] and check["outputs_equal_across_streams"]):
        raise SystemExit(f"in-flight traces differ from the synchronous trace: {check}")
    return dt, cd, check


def oracle_check(sampled, o, d, rb, outs, mb, K, who, rays_per_scene=600, strict=True):
    """The batch trace's records of the first rays_per_scene rays of each sampled sub-scene ({index: scene})
    against the CPU oracle (FP64, triangle seed, converged Newton): status, error kind, bounce count and the
    recorded patch sequence must be equal, and the hit points within 1e-10 relative."""
    import numpy as np
    from billardballraytracing3d_amd import _lib as L
    from oracle import oracle as O

    status = outs["status"].cpu().numpy(); bnc = outs["bounces"].cpu().numpy()
    ek = outs["err_kind"].cpu().numpy()
    hp = outs["hit_patch"].cpu().numpy(); hx = outs["hit_xyz"].cpu().numpy()
    rays_checked = mism = 0
    max_err = 0.0
    threads = host_threads()
    for i, sc in sampled.items():
        sl = slice(int(rb[i]), int(min(rb[i + 1], rb[i] + rays_per_scene)))
        r = O.OracleScene(sc, threads=threads).trace(o[sl], d[sl], mb, mode=L.NEWTON_CONVERGED, record_k=K,
                                                     threads=threads)
        kk = np.minimum(r.bounces, K)
        m = np.arange(K)[None, :] < kk[:, None]
        bad = (status[sl] != r.status) | (ek[sl] != r.err_kind) | (bnc[sl] != r.bounces) | \
            np.any(np.where(m, hp[sl] != r.hit_patch, False), 1)
        mism += int(bad.sum())
        rays_checked += r.status.size
        scale = np.maximum(np.linalg.norm(r.hit_xyz, axis=-1), 1.0)
        err = np.where(m, np.linalg.norm(hx[sl] - r.hit_xyz, axis=-1) / scale, 0.0)
        max_err = max(max_err, float(err.max()))
        _progress(f"{who}: oracle check scene {i}: {r.status.size} rays, {mism} mismatches so far")
    if strict and (mism or max_err > 1e-10):
        raise SystemExit(f"{who}: records differ from the CPU oracle ({mism} rays, max hit error {max_err:.3g})")
    return {"scenes": sorted(sampled), "rays": rays_checked, "rays_with_mismatch": mism,
            "max_rel_hit_error": max_err, "bounces_compared": K,
            "oracle": "CPU FP64, triangle seed, converged Newton"}


def error_counts(err_kind):
    """Rays of the last step per error kind (device int8 array)."""
    import torch
    c = torch.bincount(err_kind.to(torch.int64).clamp(0, len(ERROR_KINDS) - 1), minlength=len(ERROR_KINDS)).cpu()
    return {k: int(v) for k, v in zip(ERROR_KINDS, c.tolist()) if k != "none"}


def status_counts(status):
    import torch
    c = torch.bincount(status.to(torch.int64).clamp(0, 2), minlength=3).cpu().tolist()
    return {"escaped": int(c[0]), "trapped": int(c[1]), "error": int(c[2])}


def _progress(msg):
    print(msg, file=sys.stderr, flush=True)


def full_form_leg(args):
    """The full batched form: args.full_scenes scenes of the batched layout, args.full_rays rays in total, 200
    bounces, learned Newton seeds (network trained on the GPU on pairs of the Sinai scene) with the
    triangle-seed fallback, one hipGraph-replayed persistent launch with an autotuned shape.  Reports the
    fallback rate per scene and checks sampled scenes against the CPU oracle (triangle seed, converged)."""
    import numpy as np
    import torch
    from billardballraytracing3d_amd import scenes as S, tracer as T, _lib as L
    from oracle import oracle as O

    dev = torch.device("cuda", 0)
    t0 = time.time()
    scenes, o, d, rb = S.batch_config(args.seed, n_scenes=args.full_scenes, total_rays=args.full_rays)
    check_ids = np.linspace(0, len(scenes) - 1, 6).astype(int)
    t1 = time.time()
    _progress(f"full form: {len(scenes)} scenes, {int(rb[-1])} rays generated in {t1 - t0:.1f} s")
    gb = T.GpuScene.batch(scenes, compact=True, device_build=True)
    t2 = time.time()
    _progress(f"full form: {gb.n_patches} patches built (compact, device build) in {t2 - t1:.1f} s")
    # the network: trained on the Sinai scene (a compact handle of its own), then attached to the batch
    sinai = T.GpuScene.batch([S.sinai_scene(0)], compact=True, device_build=True)
    net = T.Mlp(args.seed + 3)
    rep, _ = net.train(sinai, n_samples=args.train_samples, seed=21 + args.seed, batch=args.train_batch,
                       epochs=args.train_epochs, lr=args.train_lr, lr_final=args.train_lr / 100)
    del sinai
    gb.set_mlp(net)
    t3 = time.time()
    _progress(f"full form: network trained in {t3 - t2:.1f} s, test loss {rep['test_loss']:.3g}")
    n = int(rb[-1])
    mb, K = 200, 8
    mode = L.NEWTON_CONVERGED_FP32 if args.full_newton == "fp32" else L.NEWTON_CONVERGED
    to = torch.from_numpy(o).to(dev)
    td = torch.from_numpy(d).to(dev)
    outs = dict(status=torch.empty(n, dtype=torch.int8, device=dev),
                err_kind=torch.empty(n, dtype=torch.int8, device=dev),
                bounces=torch.empty(n, dtype=torch.int32, device=dev),
                hit_patch=torch.empty((n, K), dtype=torch.int32, device=dev),
                hit_xyz=torch.empty((n, K, 3), dtype=torch.float64, device=dev),
                seed_stats=torch.empty((n, 2), dtype=torch.int32, device=dev))
    ptrs = {k: v.data_ptr() for k, v in outs.items()}
    st = torch.cuda.current_stream(dev).cuda_stream
    best, rows = gb.autotune(to.data_ptr(), td.data_ptr(), n, mb, ptrs, mode=mode, record_k=K, ray_begin=rb,
                             stream=st, seed=L.SEED_LEARNED)
    _progress(f"full form: autotuned {best} over {len(rows)} shapes")

    def step():
        return gb.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, mode=mode, record_k=K, ray_begin=rb,
                               stream=st, flags=L.FLAG_GRAPH, seed=L.SEED_LEARNED, **best)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    steps = max(1, min(args.steps, 5))
    ctr = np.zeros(L.NCOUNTERS, np.int64)
    ts = time.perf_counter()
    for _ in range(steps):
        c, _ = step()
        ctr += c
    torch.cuda.synchronize()
    te = time.perf_counter()
    bounces = int(ctr[L.CTR["bounces"]])
    _progress(f"full form: {steps} steps, {(te - ts) * 1e3 / steps:.1f} ms/step, {bounces / (te - ts):.4g} bounces/s"
              " (graph replay, one in flight)")
    one_in_flight = {"value": bounces / (te - ts), "ms_per_step": (te - ts) * 1e3 / steps, "steps": steps,
                     "note": "one hipGraph-replayed trace at a time"}
    in_flight_check = None
    if args.full_in_flight > 1:
        dt, cd, in_flight_check = run_in_flight(
            gb, to, td, n, mb, outs, steps, max(1, args.warmup), args.full_in_flight, ctr // steps,
            dict(mode=mode, record_k=K, ray_begin=rb, seed=L.SEED_LEARNED, flags=group_flag(args), **best))
        bounces = int(cd[:, L.CTR["bounces"]].sum())
        ts, te = 0.0, dt
        _progress(f"full form: {steps} steps, {dt * 1e3 / steps:.1f} ms/step, {bounces / dt:.4g} bounces/s"
                  f" ({args.full_in_flight} in flight)")
    # fallback rate per scene
    ss = outs["seed_stats"].cpu().numpy().astype(np.int64)
    per = []
    for i in range(len(scenes)):
        a = ss[rb[i]:rb[i + 1]].sum(axis=0)
        if a.sum() > 0:
            per.append(a[1] / a.sum())
    per = np.array(per)
    # oracle check on sampled scenes: same rays, triangle seed, converged Newton, FP64
    # with the FP32 start a classification may differ from the FP64 oracle's (measured and reported here)
    check = oracle_check({int(i): scenes[int(i)] for i in check_ids}, o, d, rb, outs, mb, K, "full form",
                         strict=(mode == L.NEWTON_CONVERGED))
    hits = int(ctr[L.CTR["seed_learned"]] + ctr[L.CTR["seed_fallback"]])
    learned_value, learned_one = bounces / (te - ts), one_in_flight["value"]
    tri = full_form_triangle(args, gb, to, td, n, mb, K, rb, outs, ptrs, st, steps, learned_value, learned_one,
                             mode) if args.full_triangle else None
    # the other precision split on the same handle (learned seed): FP64 Newton throughout, or the FP32 start
    other = L.NEWTON_CONVERGED if mode == L.NEWTON_CONVERGED_FP32 else L.NEWTON_CONVERGED_FP32
    alt = full_form_triangle(args, gb, to, td, n, mb, K, rb, outs, ptrs, st, steps, learned_value, learned_one,
                             other, seed=L.SEED_LEARNED) if args.full_other_newton else None
    if alt is not None:
        alt["newton_mode"] = "converged" if other == L.NEWTON_CONVERGED else "converged, FP32 first iterations"
        alt["note"] = "learned seed; learned_vs_triangle here is this configuration's learned value over this pass"
    return {
        "metric": METRIC, "value": bounces / (te - ts), "unit": "bounces/s", "n_gpus": 1, "steps": steps,
        "warmup": max(1, args.warmup), "ms_per_step": (te - ts) * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched-full-form-{len(scenes)}-scenes-{n}-rays", "scenes": len(scenes),
                   "patches": gb.n_patches, "rays": n, "max_bounces": mb, "record_k": K,
                   "newton_mode": "converged, FP32 first iterations" if mode == L.NEWTON_CONVERGED_FP32 else "converged",
                   "seed_mode": "learned", "scene": "compact, device build", "autotuned": best,
                   "traces_in_flight": max(1, args.full_in_flight)},
        "one_in_flight": one_in_flight, "in_flight_check": in_flight_check,
        "mean_newton_iters": {"learned": int(ctr[L.CTR["newton_iters"]]) / hits if hits else None,
                              "learned_fp32_phase": int(ctr[L.CTR["newton_iters_fp32"]]) / hits if hits else None},
        "seed_learned_per_step": int(ctr[L.CTR["seed_learned"]]) // steps,
        "seed_fallback_per_step": int(ctr[L.CTR["seed_fallback"]]) // steps,
        "fallback_rate_per_scene": {"scenes": int(per.size), "min": float(per.min()), "median": float(np.median(per)),
                                    "mean": float(per.mean()), "max": float(per.max())} if per.size else None,
        "oracle_check": check, "triangle_seed": tri, "other_newton_mode": alt,
        "rays_by_status": status_counts(outs["status"]), "errors_by_kind": error_counts(outs["err_kind"]),
        "bounces_per_step": bounces // steps, "train": {"pairs": rep["n_pairs"], "steps": rep["steps"],
                                                         "test_loss": rep["test_loss"]},
        "scene_generation_s": t1 - t0, "scene_build_s": t2 - t1, "train_s": t3 - t2,
        "counters_per_step": {k: int(ctr[v] // steps) for k, v in L.CTR.items()},
    }


def full_form_triangle(args, gb, to, td, n, mb, K, rb, outs, ptrs, st, steps, learned_value, learned_one, mode,
                       seed=None):
    """The full form's workload on the same handle, rays and bounce limit with another seed mode (default the
    triangle seed) or Newton mode, timed the same way (autotuned shape, hipGraph replays one at a time, then
    traces in flight): the learned path's ratio on this workload.  Its records are compared with the learned
    run's (the acceptance rule keeps every learned result one the triangle path gives)."""
    import numpy as np
    import torch
    from billardballraytracing3d_amd import _lib as L

    seed = L.SEED_TRIANGLE if seed is None else seed
    keep = {k: outs[k].clone() for k in ("status", "err_kind", "bounces", "hit_patch", "hit_xyz")}
    if seed == L.SEED_TRIANGLE:   # the triangle seed writes no per-ray learned/fallback counts
        outs = {k: v for k, v in outs.items() if k != "seed_stats"}
        ptrs = {k: v.data_ptr() for k, v in outs.items()}
    best, rows = gb.autotune(to.data_ptr(), td.data_ptr(), n, mb, ptrs, mode=mode, record_k=K, ray_begin=rb, stream=st,
                             seed=seed)

    def step():
        return gb.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, mode=mode, record_k=K, ray_begin=rb,
                               stream=st, flags=L.FLAG_GRAPH, seed=seed, **best)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    ctr = np.zeros(L.NCOUNTERS, np.int64)
    ts = time.perf_counter()
    for _ in range(steps):
        c, _ = step()
        ctr += c
    torch.cuda.synchronize()
    te = time.perf_counter()
    one = int(ctr[L.CTR["bounces"]]) / (te - ts)
    value, ms = one, (te - ts) * 1e3 / steps
    if args.full_in_flight > 1:
        dt, cd, _ = run_in_flight(gb, to, td, n, mb, outs, steps, max(1, args.warmup), args.full_in_flight,
                                  ctr // steps, dict(mode=mode, record_k=K, ray_begin=rb, flags=group_flag(args),
                                                     seed=seed, **best))
        value, ms = int(cd[:, L.CTR["bounces"]].sum()) / dt, dt * 1e3 / steps
    slot = torch.arange(K, device=outs["bounces"].device)[None, :]
    valid = slot < torch.clamp(keep["bounces"], max=K)[:, None].to(slot.dtype)
    same = (outs["status"] == keep["status"]) & (outs["err_kind"] == keep["err_kind"]) & \
        (outs["bounces"] == keep["bounces"]) & torch.all((outs["hit_patch"] == keep["hit_patch"]) | ~valid, dim=1)
    diff = (outs["hit_xyz"] - keep["hit_xyz"]).abs().amax(dim=2)
    max_diff = float(torch.where(valid, diff, torch.zeros_like(diff)).max().item())
    calls = int(ctr[L.CTR["newton_calls"]])
    _progress(f"full form, {'triangle' if seed == L.SEED_TRIANGLE else 'learned'} seed, newton mode {mode}: "
              f"{value:.4g} bounces/s ({args.full_in_flight} in flight), {one:.4g} one at a time")
    return {"value": value, "ms_per_step": ms, "one_in_flight": {"value": one, "ms_per_step": (te - ts) * 1e3 / steps},
            "autotuned": best, "learned_vs_triangle": learned_value / value,
            "learned_vs_triangle_one_in_flight": learned_one / one,
            "mean_newton_iters": int(ctr[L.CTR["newton_iters"]]) / calls if calls else None,
            "records_vs_learned": {"rays": n, "rays_with_mismatch": int((~same).sum().item()),
                                   "max_abs_hit_diff": max_diff, "bounces_compared": K}}


MLP_FLOPS_PER_WAVE_EVAL = 160 * 16 * 16 * 16 * 2   # 160 bf16 16x16x16 matrix products per 64-lane evaluation
BF16_PEAK_TFLOPS = 2500.0


def learned_leg(args, gs, to, td, n, mb, K, outs, ptrs, stream, flags, value_tri, res_tri):
    """Trains the seed network on the GPU (untimed), then times the same trace with the learned seed and
    checks its records against the triangle-seed run's."""
    import numpy as np
    import torch
    from billardballraytracing3d_amd import tracer as T, _lib as L

    kk = min(20, K)
    ref = {k: outs[k].clone() for k in ("status", "err_kind", "bounces")}
    if kk > 0:
        ref["hit_patch"] = outs["hit_patch"][:, :kk].clone()
        ref["hit_xyz"] = outs["hit_xyz"][:, :kk].clone()
    net = T.Mlp(args.seed + 3)
    t0 = time.perf_counter()
    rep, curve = net.train(gs, n_samples=args.train_samples, seed=21 + args.seed, batch=args.train_batch,
                           epochs=args.train_epochs, lr=args.train_lr, lr_final=args.train_lr / 100)
    t_train = time.perf_counter() - t0
    gs.set_mlp(net)

    def step():
        return gs.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, mode=L.NEWTON_CONVERGED, record_k=K,
                               flags=flags, stream=stream.cuda_stream, block=args.block,
                               waves_per_eu=args.waves_per_eu, seed=L.SEED_LEARNED, newton_batch=args.newton_batch)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctr = np.zeros(L.NCOUNTERS, np.int64)
    kms = np.zeros(3)
    ts = time.perf_counter()
    for _ in range(args.steps):
        c, k = step()
        ctr += c
        kms += k
    torch.cuda.synchronize()
    te = time.perf_counter()
    bounces = int(ctr[L.CTR["bounces"]])
    value = bounces / (te - ts)
    hits = int(ctr[L.CTR["seed_learned"]] + ctr[L.CTR["seed_fallback"]])
    evals = int(ctr[L.CTR["mlp_evals"]])
    # records against the triangle-seed run
    same = (outs["status"] == ref["status"]) & (outs["err_kind"] == ref["err_kind"]) & (outs["bounces"] == ref["bounces"])
    max_diff = 0.0
    if kk > 0:
        slot = torch.arange(kk, device=outs["bounces"].device)[None, :]
        valid = slot < torch.clamp(ref["bounces"], max=kk)[:, None].to(slot.dtype)
        same &= torch.all((outs["hit_patch"][:, :kk] == ref["hit_patch"]) | ~valid, dim=1)
        diff = (outs["hit_xyz"][:, :kk] - ref["hit_xyz"]).abs().amax(dim=2)
        max_diff = float(torch.where(valid, diff, torch.zeros_like(diff)).max().item())
    mismatch = int((~same).sum().item())
    kernel_ms = kms[0]
    mlp_tflops = evals * MLP_FLOPS_PER_WAVE_EVAL / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0

    def records_vs_triangle():
        """(rays whose records differ from the triangle-seed run's, the bounce index at which each such ray's
        record first departs: the first differing patch among the compared slots, else the bounce at which
        one of the two runs ended)."""
        nb, nb_ref = outs["bounces"], ref["bounces"]
        differ = (outs["status"] != ref["status"]) | (outs["err_kind"] != ref["err_kind"]) | (nb != nb_ref)
        ended = torch.minimum(nb, nb_ref).to(torch.int64)
        first = ended.clone()
        if kk > 0:
            slot = torch.arange(kk, device=nb.device)[None, :]
            valid = slot < torch.clamp(ended, max=kk)[:, None]
            diffp = (outs["hit_patch"][:, :kk] != ref["hit_patch"]) & valid
            anyp = diffp.any(dim=1)
            differ |= anyp
            first = torch.where(anyp, torch.argmax(diffp.to(torch.int32), dim=1).to(torch.int64), ended)
        f = first[differ].cpu().numpy()
        stats = None if f.size == 0 else {"min": int(f.min()), "median": float(np.median(f)), "max": int(f.max()),
                                          "within_20": int((f < 20).sum())}
        return int(differ.sum().item()), stats

    # the network-first form (BBRT_FLAG_SEED_FIRST): stage 1 only where the network seed's Newton is not decisive
    first = None
    if not args.no_seed_first:
        def fstep():
            return gs.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, mode=L.NEWTON_CONVERGED, record_k=K,
                                   flags=flags | L.FLAG_SEED_FIRST, stream=stream.cuda_stream, block=args.block,
                                   waves_per_eu=args.waves_per_eu, seed=L.SEED_LEARNED, newton_batch=args.newton_batch)
        for _ in range(args.warmup):
            fstep()
        torch.cuda.synchronize()
        fctr = np.zeros(L.NCOUNTERS, np.int64)
        fs = time.perf_counter()
        for _ in range(args.steps):
            c, _ = fstep()
            fctr += c
        torch.cuda.synchronize()
        fe = time.perf_counter()
        fv = int(fctr[L.CTR["bounces"]]) / (fe - fs)
        first_diff = records_vs_triangle()
        first = {"value": fv, "ms_per_step": (fe - fs) * 1e3 / args.steps,
                 "vs_triangle": fv / value_tri if value_tri else None, "vs_learned_stage1_first": fv / value,
                 "per_step": {k: int(fctr[L.CTR[k]]) // args.steps for k in
                              ("seed_learned", "seed_miss", "seed_fallback", "mt_tests", "newton_iters", "bounces")},
                 "rays_differing_from_triangle_records": first_diff[0], "first_divergent_bounce": first_diff[1],
                 "bounces_compared": kk,
                 "note": "network seed to Newton first; in-domain root = hit, converged outside the patch by more "
                         "than the margin = miss, else stage 1 + triangle seed; records measured against the "
                         "triangle-seed run (not promised equal)"}
    return {
        "value": value, "unit": "bounces/s", "ms_per_step": (te - ts) * 1e3 / args.steps,
        "vs_triangle": value / value_tri if value_tri else None,
        "mean_newton_iters": int(ctr[L.CTR["newton_iters"]]) / hits if hits else None,
        "stage1_hits_per_step": hits // args.steps,
        "seed_learned_per_step": int(ctr[L.CTR["seed_learned"]]) // args.steps,
        "seed_fallback_per_step": int(ctr[L.CTR["seed_fallback"]]) // args.steps,
        "network_evals_per_step": evals // args.steps,
        "lanes_per_eval": int(ctr[L.CTR["mlp_lanes"]]) / evals if evals else None,
        "avg_launch_ms": kernel_ms / kms[1] if kms[1] else None,
        "mlp_roofline": {"bound": "mfma", "compute": "bf16 dense", "achieved": mlp_tflops, "peak": BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": mlp_tflops / BF16_PEAK_TFLOPS},
        "check_vs_triangle": {"rays": n, "bounces_compared": kk, "rays_with_mismatch": mismatch,
                              "max_abs_hit_diff": max_diff},
        "seed_first": first,
        "train": {"pairs": rep["n_pairs"], "steps": rep["steps"], "batch": args.train_batch,
                  "epochs": args.train_epochs, "first_loss": rep["first_loss"], "final_loss": rep["final_loss"],
                  "val_loss": rep["val_loss"], "test_loss": rep["test_loss"], "gen_ms": rep["gen_ms"],
                  "train_ms": rep["train_ms"], "wall_s": t_train},
    }


def committed_profile_value(args, n, mb, K, key):
    """A per-launch figure of k_persist measured by rocprofv3 for the default workload (profiles/k_persist_sinai_1M.md:
    traffic_bytes_per_launch, fp64_valu_lane_flops_per_launch), or None for any other workload."""
    if (n, mb, K, args.mode, args.seed, args.no_cull, args.wavefront) != (1 << 20, 100, 100, "converged", 0, False, False):
        return None
    path = os.path.join(ROOT, "profiles", "k_persist_sinai_1M.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def copy_probe(mb):
    """A plain device-to-device copy of known size, for scaling the PMC byte counters."""
    import torch
    n = mb * (1 << 20) // 8
    x = torch.arange(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 3}), flush=True)


def host_threads():
    """Host threads for the CPU legs: OMP_NUM_THREADS when set (the process's share of the host's cores),
    else every core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "").strip()
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def host_cpu():
    """nproc, the cores this process may use and the CPU model, as recorded beside the CPU baseline."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "model": model}


def cpu_baseline(sc, o, d, mb, budget_s, gs, K):
    """The CPU oracle timed as the reference runs: brute-force stage 1, the reference's Newton stopping rule
    (step < 1e-3, SurfacePatch.hpp:629), FP64, OpenMP dynamic over rays (PlaneSourceDriver.hpp:96), on all the
    host threads this process may use, over consecutive 2,048-ray chunks of the same batch until the time budget
    is spent.  The GPU's records for those rays, traced in the same Newton mode, are checked against it."""
    import numpy as np
    from billardballraytracing3d_amd import _lib as L
    from oracle import oracle as O

    mode = L.NEWTON_REF
    threads = host_threads()
    osc = O.OracleScene(sc, threads=threads)
    chunk = 2048
    t_tot = 0.0
    bounces = 0
    rays = 0
    max_err = 0.0
    seq_mismatch = 0
    start = 0
    while t_tot < budget_s and start < o.shape[0]:
        sl = slice(start, start + chunk)
        t0 = time.perf_counter()
        r = osc.trace(o[sl], d[sl], mb, mode=mode, threads=threads)
        t_tot += time.perf_counter() - t0
        bounces += int(r.bounces.sum())
        rays += r.bounces.shape[0]
        g = gs.trace(o[sl], d[sl], mb, mode=mode, record_k=mb)
        kk = min(20, mb)
        seq_mismatch += int(np.sum(np.any(g.hit_patch[:, :kk] != r.hit_patch[:, :kk], axis=1) |
                                   (g.status != r.status) | (g.err_kind != r.err_kind) | (g.bounces != r.bounces)))
        scale = np.maximum(np.linalg.norm(r.hit_xyz, axis=-1), 1.0)
        max_err = max(max_err, float(np.max(np.linalg.norm(g.hit_xyz - r.hit_xyz, axis=-1) / scale)))
        start += chunk
    return {"value": bounces / t_tot if t_tot > 0 else None, "unit": "bounces/s", "cores": threads, "kind": "port",
            "sample": f"first {rays} rays of the same batch ({bounces} bounces) in {t_tot:.1f} s, oracle "
                      f"brute-force stage 1, reference Newton rule (step < 1e-3), OpenMP dynamic",
            "host": host_cpu(),
            "gpu_check": {"rays": rays, "newton_mode": "ref", "max_rel_hit_error": max_err,
                          "rays_with_sequence_mismatch": seq_mismatch}}


if __name__ == "__main__":
    main()
