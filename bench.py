"""Benchmark: closest-hit Mrays/s on 2^24 incoherent rays over the 8.39M-triangle
heightfield terrain (2048x2048 cells), SAH KD tree built on the CPU and
uploaded once, one MI355X; and the equal-RMSE wall time of the radiance-cache
path tracer against the plain path tracer on the 100k-triangle Cornell room.

A step is one closest-hit launch over the whole ray stream (rays resident in
HBM).  Prints ONE JSON line with the metric, a roofline object for the
traversal kernel, the radiance-cache comparison, the batched multi-scene
render (64 scenes in one wavefront, key "batch"), the ray tracer on the
25,602-triangle spheres scene (key "ray_tracer") and a CPU baseline (the
restated reference CPU path timed on this host).
Run:  python bench.py [--steps K] [--warmup W]

Other workloads, one JSON line each: --workload batch (64 scenes path traced
in one wavefront), --workload io (flat tree files, progressive path tracing,
the drop-in CLI on the example scenes), --workload cache (the radiance-cache
comparison alone), --workload build (the KD-tree builds on the GPU),
--workload rt (the ray tracer on the spheres scene).
"""
import argparse
import json
import os

# OpenMP placement of the CPU baseline (one thread per core, packed), set
# before anything loads an OpenMP runtime: libgomp reads these once, when the
# host library that builds the trees is loaded.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# With a binding set, libgomp pins the loading thread to its first place, so
# the CPUs this process may use are counted before that happens.
ALL_CPUS = os.sched_getaffinity(0)
USABLE_CPUS = len(ALL_CPUS)
OMP_MASTER_CPUS = ALL_CPUS  # the loading thread's place, once libgomp has pinned it

import platform  # noqa: E402
import subprocess  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "Mrays/s closest-hit (incoherent, 8M tris); equal-RMSE time, cache vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    """The host CPU as lscpu describes it, with the CPUs this process may run
    on: {model, vendor, sockets, cores_per_socket, threads_per_core,
    physical_cores, logical_cpus, usable_cpus}."""
    info = {}
    keys = {"Model name": "model", "Vendor ID": "vendor", "Socket(s)": "sockets",
            "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "threads_per_core",
            "CPU(s)": "logical_cpus", "CPU family": "family", "Model": "model_number"}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in keys and keys[k] not in info:  # first occurrence: the whole machine's
                info[keys[k]] = v.strip()
    except Exception:
        pass
    info.setdefault("model", platform.processor() or "unknown")
    for k in ("sockets", "cores_per_socket", "threads_per_core", "logical_cpus"):
        try:
            info[k] = int(info[k])
        except (KeyError, ValueError):
            info.pop(k, None)
    if "sockets" in info and "cores_per_socket" in info:
        info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    info["usable_cpus"] = USABLE_CPUS
    return info


def cpu_model():
    return cpu_info()["model"]


def load_libraries(P):
    """Loads the native libraries.  libgomp, loaded with the host library,
    pins this thread to its first place (OMP_PROC_BIND); the thread's full
    CPU set is then restored, so that the threads the GPU runtime starts later
    and the processes this run spawns (the CLIs, which inherit it) are not
    confined to that one core.  The CPU baseline re-pins the thread to its
    place while it runs (omp_master_pinned)."""
    global OMP_MASTER_CPUS
    P.libraries()
    OMP_MASTER_CPUS = os.sched_getaffinity(0)
    os.sched_setaffinity(0, ALL_CPUS)


class omp_master_pinned:
    """This thread on its OpenMP place (the first core) for the duration."""

    def __enter__(self):
        os.sched_setaffinity(0, OMP_MASTER_CPUS)

    def __exit__(self, *exc):
        os.sched_setaffinity(0, ALL_CPUS)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--det-eps", type=float, default=1e-12,
                    help="determinant cull; the reference's 1e-4 culls every terrain triangle")
    ap.add_argument("--layout", type=int, default=-1)
    ap.add_argument("--max-depth", type=int, default=40,
                    help="KD-tree depth cap of the trees built with this build's parameters (reference: 20)")
    ap.add_argument("--leaf-size", type=int, default=8, help="nodes this small are not split (reference: 32)")
    ap.add_argument("--traversal-cost", type=float, default=2.0, help="SAH traversal cost (reference: 5)")
    ap.add_argument("--no-autotune", action="store_true", help="trace with the default tuning")
    ap.add_argument("--tuning", default="",
                    help="explicit traversal tuning instead of the autotuner: comma-separated values of "
                         "layout,stack_lds,leaf_ilp,fetch_min,occupancy,wg_waves,leaf_batch,node_levels "
                         "(a prefix of them)")
    ap.add_argument("--no-ref-tree", action="store_true",
                    help="skip timing the same rays over the tree built with the reference's parameters (when the "
                         "timed tree is another)")
    ap.add_argument("--tree", choices=("ref-sah", "cpu-sah", "cpu-sah-events", "gpu-sah"), default="ref-sah",
                    help="the timed tree: the reference's SAH KD tree (its event sweep and parameters: depth cap 20, "
                         "leaf size 32, costs 5 / 1.5; default), or one built with --max-depth, --leaf-size and "
                         "--traversal-cost on the CPU with 64 bins per axis or the event sweep, all uploaded (the "
                         "workload's definition), or on the GPU")
    ap.add_argument("--no-other-tree", action="store_true",
                    help="skip timing the same rays over the trees built the other ways")
    ap.add_argument("--no-ref-eps", action="store_true",
                    help="skip timing the same rays with the reference's det_eps 1e-4")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rays", type=int, default=1 << 24, help="CPU baseline sample (prefix of the stream)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run one plain device copy of this many MiB (reference for HBM counters)")
    ap.add_argument("--no-cache", action="store_true", help="skip the radiance-cache equal-RMSE measurement")
    ap.add_argument("--cache-size", type=int, default=1024, help="image side of the radiance-cache measurement")
    ap.add_argument("--ref-spp", type=int, default=4096, help="samples per pixel of the plain reference image")
    ap.add_argument("--workload", choices=("terrain", "batch", "io", "cache", "build", "pt", "rt"), default="terrain",
                    help="terrain: the closest-hit stream (default); batch: 64 scenes path traced in one wavefront; "
                         "io: flat tree files, progressive path tracing and the CLI; cache: the radiance-cache "
                         "equal-RMSE measurement alone; build: the KD-tree builds on the GPU; pt: one plain and "
                         "one radiance-cache render of the Cornell room (--cache-size, --pt-spp, --pt-frames); rt: "
                         "the ray tracer on the 25,602-triangle spheres scene at 512x512 with both CLIs")
    ap.add_argument("--pt-spp", type=int, default=16, help="samples per pixel of the pt workload's plain render")
    ap.add_argument("--pt-frames", type=int, default=16, help="frames of the pt workload's cache render")
    ap.add_argument("--cache-spf", type=int, default=8,
                    help="samples per pixel in each radiance-cache frame (one training step per frame)")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-scene leg of the default run")
    ap.add_argument("--no-rt", action="store_true", help="skip the ray-tracer leg (spheres scene) of the default run")
    ap.add_argument("--rt-tuning", default="", help="traversal tuning of the ray-tracer leg (as --tuning)")
    ap.add_argument("--path-budget", type=int, default=-1,
                    help="paths the plain path tracer keeps in flight (0: one sample per pass; default: the "
                         "library's, 2^24)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="path tracers enqueue their passes kernel by kernel instead of replaying hipGraphs")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the ray stream would shard by ray index across GPUs (each "
                                   "holding the whole scene), but --gpus > 1 is not supported"}))
        sys.exit(2)

    import nostrum_pray_amd as P

    load_libraries(P)
    if args.workload == "batch":
        batch_workload(P, args)
        return
    if args.workload == "io":
        io_workload(P, args)
        return
    if args.workload == "build":
        build_workload(P, args)
        return
    if args.workload == "pt":
        pt_workload(P, args)
        return
    if args.workload == "rt":
        rt_workload(P, args)
        return
    if args.workload == "cache":
        ctx = P.Context(0)
        configure(ctx, args)
        try:
            out = radiance_cache_leg(P, ctx, args)
        finally:
            ctx.close()
        print(json.dumps({"metric": "equal-RMSE time, radiance cache vs plain path tracing", "n_gpus": 1,
                          "data": "synthetic", "radiance_cache": out}), flush=True)
        return

    # The ray tracer first, in a context of its own and before this process
    # holds anything else, as a ray-tracing application's process would:
    # measured after the terrain leg (its context, 8.4M-triangle scene and
    # 2^24-ray buffers still held), the same frames took 0.19 ms longer on
    # the host side (kernel times unchanged).
    rt = None
    if not args.no_rt:
        rctx = P.Context(0)
        configure(rctx, args)
        rt = ray_tracer_leg(P, rctx, args)
        rctx.close()

    t0 = time.time()
    scene = P.Scene.terrain(args.grid, args.seed)
    n_tris = scene.num_tris
    ctx = P.Context(0)
    configure(ctx, args)
    sid, tree_tag, tree_desc, tree_build, (tree_depth, tree_nodes, tree_refs), raw_box = \
        build_tree(P, ctx, scene, args, args.tree)
    log(f"scene {n_tris} tris, tree {tree_tag}: depth {tree_depth}, {tree_nodes} nodes, {tree_refs} references, "
        f"{tree_build} ({time.time() - t0:.1f}s)")

    tuning = None
    if args.tuning:
        ctx.set_tuning(sid, **parse_tuning(P, args.tuning))
    elif args.layout >= 0:
        ctx.set_tuning(sid, layout=args.layout, node_levels=1 if args.layout == P.LAYOUT_INDEXED else -1)
    elif not args.no_autotune:
        t1 = time.time()
        tuning = ctx.autotune(sid, n_rays=n_rays_for_tuning(args), force=True)
        log(f"autotuned {tuning} in {time.time() - t1:.1f}s")
    box = P.scene_box_expanded(raw_box)

    # Inputs and outputs resident in HBM for the whole run.
    n = args.rays
    d_rays = ctx.alloc((n, 6), np.float32)
    d_idx = ctx.alloc(n, np.int32)
    d_t = ctx.alloc(n, np.float32)
    d_cnt = ctx.alloc((n, 2), np.uint32)
    d_lim = ctx.to_device(np.full(n, np.finfo(np.float32).max, np.float32))
    ctx.gen_rays_device(args.seed, box, n, d_rays.ptr)
    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        src, dst = ctx.alloc(nb, np.uint8), ctx.alloc(nb, np.uint8)
        ctx.copy_device(dst.ptr, src.ptr, nb)
        ctx.sync()
        log(f"calibration copy: {nb} bytes read, {nb} bytes written")
        src.free()
        dst.free()

    # Work counts (untimed): nodes and leaf references per ray, of the
    # uploaded KD tree's traversal (plane-form node steps) and of the node
    # form the tuning runs.
    tun = ctx.get_tuning(sid)

    def work_counts(mode, lim=None, plane_form=False):
        if plane_form:
            ctx.set_tuning(sid, node_levels=1)
        ctx.trace_counts_device(sid, d_rays.ptr, n, d_idx.ptr, d_t.ptr, d_cnt.ptr, mode, args.det_eps,
                                d_limit2=lim)
        ctx.sync()
        if plane_form:
            ctx.set_tuning(sid, **tun)
        return d_cnt.numpy().astype(np.int64).sum(0) / n

    kd_cnt = work_counts(P.TRACE_CLOSEST, plane_form=True)
    hits = int((d_idx.numpy() >= 0).sum())
    own_cnt = work_counts(P.TRACE_CLOSEST)
    d_cnt2 = d_cnt  # reused for the other tree's counts
    # Algorithmic bytes per ray of the traversal that runs: ray 24 + result
    # 8 + the bytes of each node visited + 40 per triangle reference tested
    # (4-byte index + a, ab, ac; the packed layout moves 48 bytes per
    # reference).  A KD node is 8 bytes; the boxed node forms (node records 3
    # and 4) run over a finer tree of their own whose children cost 28 bytes
    # each (box and word).  The uploaded KD tree's own work per ray is
    # reported beside.
    boxed = tun["node_levels"] in (3, 4)
    nodes_per_ray, refs_per_ray = own_cnt
    bytes_per_ray = boxed_bytes(own_cnt) if boxed else kd_bytes(own_cnt)
    log(f"hits {hits}/{n}  {'children' if boxed else 'nodes'}/ray {nodes_per_ray:.1f}  refs/ray {refs_per_ray:.1f}  "
        f"B/ray {bytes_per_ray:.0f}  (KD tree: nodes/ray {kd_cnt[0]:.1f} refs/ray {kd_cnt[1]:.1f} "
        f"B/ray {kd_bytes(kd_cnt):.0f})")

    def timed(mode, steps, warmup, lim=None, scene_id=sid, eps=args.det_eps):
        for _ in range(warmup):
            ctx.trace_device(scene_id, d_rays.ptr, n, d_idx.ptr, d_t.ptr, mode, eps, d_limit2=lim)
        ctx.sync()
        t_wall = time.perf_counter()
        ctx.record(0)
        for _ in range(steps):
            ctx.trace_device(scene_id, d_rays.ptr, n, d_idx.ptr, d_t.ptr, mode, eps, d_limit2=lim)
        ctx.record(1)
        ms_ev = ctx.elapsed_ms(0, 1) / steps
        ctx.sync()
        return ms_ev, (time.perf_counter() - t_wall) * 1e3 / steps

    # Timed region: HIP events on the stream the kernels run on.
    ms, ms_wall = timed(P.TRACE_CLOSEST, args.steps, args.warmup)
    mrays = n / (ms * 1e-3) / 1e6
    log(f"closest-hit {ms:.3f} ms/step  {mrays:.1f} Mrays/s  (wall {ms_wall:.3f} ms/step)")
    closest_idx, closest_t = d_idx.numpy(), d_t.numpy()

    # The roofline's kernel time is the timed region's own: HIP events on the
    # kernel's stream around the K back-to-back launches, divided by K (each
    # launch is the traversal kernel and a one-wave reset of its 64 ray
    # counters, about 3 us, in front of it).  Beside it, events around each
    # traversal kernel alone, enqueued back to back and read once at the end:
    # kernels timed that way run 1-2% longer in a profiler's kernel trace than
    # those of the timed region, so the region is the figure used.
    def kernel_time(mode, lim=None):
        ctx.kernel_timing(True)
        for _ in range(args.steps):
            ctx.trace_device(sid, d_rays.ptr, n, d_idx.ptr, d_t.ptr, mode, args.det_eps, d_limit2=lim)
        kms, cnt = ctx.timed_kernels_ms()
        ctx.kernel_timing(False)
        assert cnt == min(args.steps, 64), cnt
        return float(kms)

    kernel_ms = ms
    isolated_ms = kernel_time(P.TRACE_CLOSEST)
    log(f"traversal kernel {kernel_ms:.3f} ms/launch (events around each kernel alone {isolated_ms:.3f} ms)")
    achieved = bytes_per_ray * n / (kernel_ms * 1e-3) / 1e9

    # Any-hit over the same stream (unlimited distance), reported beside with
    # its own work counts and kernel time.
    ms_any, _ = timed(P.TRACE_ANY, args.steps, 2, lim=d_lim.ptr)
    acnt = work_counts(P.TRACE_ANY, d_lim.ptr, plane_form=True)
    any_kd_bytes = kd_bytes(acnt)
    if boxed:
        acnt = work_counts(P.TRACE_ANY, d_lim.ptr)
    any_bytes = boxed_bytes(acnt) if boxed else kd_bytes(acnt)
    any_kernel_ms = ms_any
    any_isolated_ms = kernel_time(P.TRACE_ANY, d_lim.ptr)
    any_achieved = any_bytes * n / (any_kernel_ms * 1e-3) / 1e9
    log(f"any-hit {ms_any:.3f} ms/step  {n / (ms_any * 1e-3) / 1e6:.1f} Mrays/s  kernel {any_kernel_ms:.3f} ms  "
        f"nodes/ray {acnt[0]:.1f}  refs/ray {acnt[1]:.1f}  {any_achieved:.0f} GB/s")

    # The reference's determinant cull (1e-4) over the same stream: below the
    # terrain triangles' scale, so nothing is hit.
    ref_eps = None
    if not args.no_ref_eps:
        ms_ref_eps, _ = timed(P.TRACE_CLOSEST, max(3, args.steps // 4), 1, eps=1e-4)
        ref_eps = {"det_eps": 1e-4, "hit_fraction": round(float((d_idx.numpy() >= 0).mean()), 4),
                   "ms_per_step": round(ms_ref_eps, 4), "Mrays_per_s": round(n / (ms_ref_eps * 1e-3) / 1e6, 2)}
        log(f"det_eps 1e-4: {ref_eps}")
    # HBM traffic of this kernel on this workload from the committed
    # rocprofv3 PMC summary (FETCH_SIZE/WRITE_SIZE calibrated on a plain copy)
    traffic = pmc_traffic(args, tree_tag, tun) if args.rays == 1 << 24 else None

    # The same rays over the SAH trees built the other ways (each with its
    # own autotuned setting): identical hits.
    others = {}
    if not args.no_other_tree:
        for which in TREES:
            if which in (args.tree, "ref-sah"):  # the reference's tree has a leg of its own
                continue
            osid, _, odesc, obuild, (od, on, orf), _ = build_tree(P, ctx, scene, args, which)
            if tuning is not None:
                ctx.autotune(osid, n_rays=n_rays_for_tuning(args), force=True)
            else:
                ctx.set_tuning(osid, **tun)
            otun = ctx.get_tuning(osid)
            ctx.trace_counts_device(osid, d_rays.ptr, n, d_idx.ptr, d_t.ptr, d_cnt2.ptr, P.TRACE_CLOSEST,
                                    args.det_eps)
            ocnt = d_cnt2.numpy().astype(np.int64).sum(0) / n
            ms_o, _ = timed(P.TRACE_CLOSEST, max(3, args.steps // 4), 1, scene_id=osid)
            same_o = bool(np.array_equal(closest_idx, d_idx.numpy()) and
                          np.array_equal(closest_t.view(np.uint32), d_t.numpy().view(np.uint32)))
            ctx.free_scene(osid)
            others[which] = {"tree": f"{odesc}: depth {od}, {on} nodes, {orf} references", **obuild,
                             "tuning": otun, "nodes_per_ray": round(float(ocnt[0]), 2),
                             "refs_per_ray": round(float(ocnt[1]), 2), "ms_per_step": round(ms_o, 4),
                             "Mrays_per_s": round(n / (ms_o * 1e-3) / 1e6, 2), "identical_hits": same_o}
            log(f"{which}: {odesc}, {obuild}: {ms_o:.3f} ms/step = {n / (ms_o * 1e-3) / 1e6:.1f} Mrays/s, "
                f"identical hits {same_o}")

    # The same rays over the tree built with the reference's parameters
    # (depth cap 20, leaf size 32, SAH costs 5 / 1.5), when the timed tree is
    # another: identical hits.
    ref_tree = None
    if not args.no_ref_tree and args.tree != "ref-sah":
        t1 = time.time()
        rtree = scene.build(1)
        rdesc = TREES["ref-sah"]
        rsid = ctx.upload(rtree, shading=False)
        if tuning is not None:  # its own setting: its 32-triangle leaves favour other refill thresholds
            ctx.autotune(rsid, n_rays=n_rays_for_tuning(args), force=True)
        else:
            ctx.set_tuning(rsid, **tun)
        log(f"reference-parameter tree {rtree.stats()} in {time.time() - t1:.1f}s")
        ms_ref, _ = timed(P.TRACE_CLOSEST, max(3, args.steps // 4), 1, scene_id=rsid)
        same = bool(np.array_equal(closest_idx, d_idx.numpy()) and
                    np.array_equal(closest_t.view(np.uint32), d_t.numpy().view(np.uint32)))
        rtun = ctx.get_tuning(rsid)
        ctx.free_scene(rsid)
        ref_tree = {"ms_per_step": round(ms_ref, 4), "Mrays_per_s": round(n / (ms_ref * 1e-3) / 1e6, 2),
                    "identical_hits": same, "tree": rdesc, "tuning": rtun}
        log(f"reference-parameter tree: {ms_ref:.3f} ms/step, identical hits {same}")

    rc_out = None
    if not args.no_cache:
        for b in (d_rays, d_idx, d_t, d_lim, d_cnt2):
            b.free()
        d_rays = d_idx = d_t = d_lim = d_cnt2 = None
        ctx.free_scene(sid)
        sid = None
        rc_out = radiance_cache_leg(P, ctx, args)

    batch = None
    if not args.no_batch:
        batch = batch_leg(P, ctx, args)
        batch.pop("autotuned_buckets", None)
        # per scene: [triangles, image side, rays, Mrays/s rendered alone, image equal to the batch's]
        batch["per_scene"] = [[p["triangles"], p["image"], p["rays"], p["Mrays_per_s_alone"],
                               p["Msegments_per_s_alone"], p["equal_to_batch"]] for p in batch["per_scene"]]
        batch["per_scene_fields"] = ["triangles", "image_side", "rays", "Mrays_per_s_alone", "Msegments_per_s_alone",
                                     "equal_to_batch"]


    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, args, box, closest_idx)

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"closest-hit, {n} incoherent rays (origins uniform in the scene box +10%, directions "
                        f"uniform on the sphere, Philox seed {args.seed}), heightfield terrain {args.grid}x{args.grid} "
                        f"cells = {n_tris} FP32 triangles, {tree_desc}, det_eps {args.det_eps:g}",
            "rays": n, "triangles": n_tris, "tree": tree_tag, "tree_depth": tree_depth,
            "tree_nodes": tree_nodes, "leaf_refs": tree_refs, "tree_build": tree_build,
            "layout": "packed" if tun["layout"] == P.LAYOUT_PACKED else "indexed",
            "tuning": tun, "autotuned": tuning is not None,
            "hit_fraction": round(hits / n, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            **({"traffic_GBs": round(traffic / (kernel_ms * 1e-3) / 1e9, 1),
                "traffic_frac": round(traffic / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)} if traffic else {}),
            "kernel": {2: "k_trace_q (closest hit, two tree levels per node step)",
                       3: "k_trace (closest hit, boxed children pairs)",
                       4: "k_trace (closest hit, 4-wide boxed records)"}.get(tun["node_levels"],
                                                                            "k_trace (closest hit)"),
            "kernel_ms": round(kernel_ms, 4),
            "kernel_ms_events_per_kernel": round(isolated_ms, 4),
            "bytes_per_ray": round(float(bytes_per_ray), 1),
            "bytes_basis": ("24 + 8 + 28 per boxed child tested + 40 per reference" if boxed else
                            "24 + 8 + 8 per node + 40 per reference"),
            ("children_per_ray" if boxed else "nodes_per_ray"): round(float(nodes_per_ray), 2),
            "refs_per_ray": round(float(refs_per_ray), 2),
            "kd_tree_traversal": {"nodes_per_ray": round(float(kd_cnt[0]), 2),
                                  "refs_per_ray": round(float(kd_cnt[1]), 2),
                                  "bytes_per_ray": round(kd_bytes(kd_cnt), 1)},
        },
        "any_hit": {"ms_per_step": round(ms_any, 4), "Mrays_per_s": round(n / (ms_any * 1e-3) / 1e6, 2),
                    "kernel_ms": round(any_kernel_ms, 4), "kernel_ms_events_per_kernel": round(any_isolated_ms, 4),
                    "bytes_per_ray": round(float(any_bytes), 1),
                    ("children_per_ray" if boxed else "nodes_per_ray"): round(float(acnt[0]), 2),
                    "refs_per_ray": round(float(acnt[1]), 2),
                    "achieved_GBs": round(any_achieved, 1), "frac": round(any_achieved / HBM_PEAK_GBS, 4),
                    "kd_tree_bytes_per_ray": round(any_kd_bytes, 1)},
        "reference_det_eps": ref_eps,
        "reference_parameter_tree": ref_tree,
        "other_trees": others,
        "radiance_cache": rc_out,
        "batch": batch,
        "ray_tracer": rt,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    for b in (d_rays, d_idx, d_t, d_lim, d_cnt2):
        if b is not None:
            b.free()
    if sid is not None:
        ctx.free_scene(sid)
    ctx.close()


def batch_leg(P, ctx, args):
    """Random icosphere/box/terrain scenes with triangle counts log-uniform in
    [1k, 2M] and images of 128^2 .. 1024^2, 16 spp, depth 6, trees built on the
    GPU: autotuned per size bucket, path traced together in one wavefront and
    each alone (the two must agree bit for bit).  Returns the aggregate and
    per-scene figures."""
    rng = np.random.default_rng(args.seed)
    n = args.batch_scenes
    tris = np.exp(rng.uniform(np.log(1e3), np.log(2e6), n)).astype(np.int64)
    sides = rng.choice([128, 256, 512, 1024], n)
    t0 = time.time()
    scenes, sids, build_ms = [], [], []
    for k in range(n):
        s = P.Scene.random(int(args.seed * 1000 + k), int(tris[k]))
        sid, bms = gpu_sah_scene(P, ctx, s, args)
        sids.append(sid)
        build_ms.append(bms)
        scenes.append(s)
    build_s = time.time() - t0
    n_tris = int(sum(s.num_tris for s in scenes))
    log(f"batch: {n} scenes ({n_tris} triangles) generated, built on the GPU and uploaded in {build_s:.1f}s "
        f"(device build time {sum(build_ms):.1f} ms)")
    tuned = {}
    for k in range(n):
        tun = ctx.autotune(sids[k])
        b = int(np.floor(np.log(scenes[k].num_tris) / np.log(4)))
        tuned.setdefault(b, tun)
    cams = [P_camera(P, s, int(sides[k])) for k, s in enumerate(scenes)]
    sizes = [(int(v), int(v)) for v in sides]
    seeds = [17 + k for k in range(n)]
    bgs = [s.info()["background"] for s in scenes]
    spp, depth, rr = 16, 6, 3
    # The batch runs with the setting of the largest image in the size bucket
    # holding most pixels: tune that scene on its own path-tracing render.
    bucket = [int(np.floor(np.log(s.num_tris) / np.log(4))) for s in scenes]
    pix = {}
    for k in range(n):
        pix[bucket[k]] = pix.get(bucket[k], 0) + sizes[k][0] * sizes[k][1]
    main = max(pix, key=pix.get)
    rep = max((k for k in range(n) if bucket[k] == main), key=lambda k: sizes[k][0] * sizes[k][1])
    batch_tuning = ctx.autotune_pt(sids[rep], cams[rep], sizes[rep][0], sizes[rep][1], depth, rr, seeds[rep],
                                   bgs[rep], 1e-9)
    log(f"batch: tuned on scene {rep} ({scenes[rep].num_tris} triangles, {sizes[rep][0]}^2): {batch_tuning}")
    imgs, _ = ctx.render_pt_batch(sids, cams, sizes, 1, depth, rr, seeds, bgs, 1e-9)  # warm-up
    imgs, ms = ctx.render_pt_batch(sids, cams, sizes, spp, depth, rr, seeds, bgs, 1e-9)
    rays = ctx.last_render_rays()
    # Path segments: the rays a tracer that traces every sample's camera ray
    # itself would trace (the renderers trace each pixel's camera ray once
    # per call and share its hit among the samples).
    n_pix = sum(w * h for w, h in sizes)
    segs = rays + (spp - 1) * n_pix
    log(f"batch: {ms:.1f} ms, {rays} rays traced ({rays / ms / 1e3:.1f} Mrays/s), {segs} path segments "
        f"({segs / ms / 1e3:.1f} M/s)")
    per, same, single_ms = [], True, 0.0
    for k in range(n):
        ctx.record(0)
        hdr, _ = ctx.render_pt(sids[k], cams[k], sizes[k][0], sizes[k][1], spp, depth, rr, seeds[k], bgs[k], 1e-9,
                               want_rgb=False)
        ctx.record(1)
        kms = ctx.elapsed_ms(0, 1)
        kr = ctx.last_render_rays()
        ok = bool(np.array_equal(hdr.view(np.uint32), imgs[k].view(np.uint32)))
        same = same and ok
        single_ms += kms
        ks = kr + (spp - 1) * sizes[k][0] * sizes[k][1]
        per.append({"triangles": scenes[k].num_tris, "image": sizes[k][0], "ms_alone": round(kms, 3),
                    "rays": kr, "Mrays_per_s_alone": round(kr / kms / 1e3, 2),
                    "Msegments_per_s_alone": round(ks / kms / 1e3, 2), "equal_to_batch": ok})
    for sid in sids:
        ctx.free_scene(sid)
    return {
        "workload": f"{n} random icosphere/box/terrain scenes, triangles log-uniform in [1k, 2M], images uniform "
                    f"over 128^2..1024^2, {spp} spp, depth {depth}, roulette from bounce {rr}, one wavefront for all "
                    f"scenes, SAH KD trees built on the GPU",
        "scenes": n, "triangles": n_tris, "spp": spp, "max_depth": depth,
        "build_wall_s": round(build_s, 2), "gpu_build_ms_total": round(float(sum(build_ms)), 2),
        "batch_ms": round(ms, 2), "rays": rays, "Mrays_per_s": round(rays / ms / 1e3, 2),
        "path_segments": segs, "Msegments_per_s": round(segs / ms / 1e3, 2),
        "rays_note": "rays traced, each pixel's camera ray once per render call (its hit is shared by the pixel's "
                     "samples); path segments also count the camera ray of every sample",
        "sum_of_single_scene_ms": round(single_ms, 2),
        "batch_speedup_over_single_scene_renders": round(single_ms / ms, 3),
        "all_scenes_equal_to_single_renders": same,
        "autotuned_buckets": {str(k): v for k, v in sorted(tuned.items())},
        "batch_tuning": batch_tuning,
        "per_scene": per,
    }


def batch_workload(P, args):
    """The batched multi-scene render alone, as one JSON line."""
    ctx = P.Context(0)
    configure(ctx, args)
    b = batch_leg(P, ctx, args)
    ctx.close()
    out = {
        "metric": "Mrays/s path tracing, batched scenes (16 spp, depth 6)",
        "value": b["Mrays_per_s"], "unit": "Mrays/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": b["workload"], "scenes": b["scenes"]},
        **{k: v for k, v in b.items() if k not in ("workload",)},
    }
    print(json.dumps(out), flush=True)


TREES = {
    "ref-sah": "the reference's SAH KD tree built on the CPU (its event sweep and parameters: depth cap 20, leaf "
               "size 32, costs 5 / 1.5)",
    "cpu-sah": "SAH KD tree built on the CPU (64 bins per axis)",
    "cpu-sah-events": "SAH KD tree built on the CPU (the reference's event sweep)",
    "gpu-sah": "SAH KD tree built on the GPU (64 bins per axis)",
}


def io_workload(P, args):
    """The flat tree file and progressive path tracing, timed: the terrain's
    binned SAH tree built on the CPU, written to a flat tree file, read back
    and uploaded (traces over both uploads must agree); then the Cornell room
    at --cache-size^2 path traced 16 spp in one call and in 16 progressive
    calls of one sample each (the images must agree bit for bit), with the
    accumulation file and the PFM written and read back.  One JSON line."""
    import shutil
    import tempfile

    ctx = P.Context(0)
    configure(ctx, args)
    tmp = tempfile.mkdtemp(prefix="pray_io_")
    try:
        scene = P.Scene.terrain(args.grid, args.seed)
        t0 = time.perf_counter()
        tree = scene.build(2, max_depth=args.max_depth, leaf_size=args.leaf_size,
                           traversal_cost=args.traversal_cost)
        build_s = time.perf_counter() - t0
        path = os.path.join(tmp, "terrain.pflat")
        t0 = time.perf_counter()
        tree.save(path)
        save_s = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        loaded = P.Tree.load(path)
        load_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        sid_l = ctx.upload(loaded, shading=False)
        ctx.sync()
        upload_s = time.perf_counter() - t0
        sid_b = ctx.upload(tree, shading=False)
        n = 1 << 20
        rays = ctx.gen_rays(args.seed, P.scene_box_expanded(tree.box()), n)
        ib, tb = ctx.trace(sid_b, rays, P.TRACE_CLOSEST, args.det_eps)
        il, tl = ctx.trace(sid_l, rays, P.TRACE_CLOSEST, args.det_eps)
        same_tree = bool(np.array_equal(ib, il) and np.array_equal(tb.view(np.uint32), tl.view(np.uint32)))
        ctx.free_scene(sid_b)
        ctx.free_scene(sid_l)
        del tree, loaded
        log(f"io: build {build_s:.2f}s, save {save_s:.2f}s ({size / 1e6:.0f} MB), load {load_s:.2f}s, "
            f"upload {upload_s:.2f}s, identical traces {same_tree}")

        w = h = args.cache_size
        room = P.Scene.cornell(1, 100_000)
        rtree = room.build(2, max_depth=args.max_depth, leaf_size=args.leaf_size, traversal_cost=args.traversal_cost)
        sid = ctx.upload(rtree)
        inf = room.info()
        cam, bg = room.camera(), inf["background"]
        spp, depth, rr, seed = 16, 8, 3, 7
        ctx.render_pt(sid, cam, w, h, 1, depth, rr, seed, bg, 1e-9, want_rgb=False)  # warm-up
        t0 = time.perf_counter()
        one, _ = ctx.render_pt(sid, cam, w, h, spp, depth, rr, seed, bg, 1e-9, want_rgb=False)
        one_ms = (time.perf_counter() - t0) * 1e3
        acc = None
        t0 = time.perf_counter()
        for s in range(spp):
            hdr, _, acc = ctx.render_pt_progressive(sid, cam, w, h, s, 1, depth, rr, seed, bg, 1e-9, accum=acc,
                                                    want_rgb=False)
        prog_ms = (time.perf_counter() - t0) * 1e3
        same_img = bool(np.array_equal(one.view(np.uint32), hdr.view(np.uint32)))
        fp = room.fingerprint()
        apath, ppath = os.path.join(tmp, "state.acc"), os.path.join(tmp, "image.pfm")
        t0 = time.perf_counter()
        P.save_accum(apath, acc, spp, fp)
        acc_save_ms = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        acc2, n2, fp2 = P.load_accum(apath)
        acc_load_ms = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        P.write_pfm(ppath, hdr)
        pfm_ms = (time.perf_counter() - t0) * 1e3
        same_files = bool(np.array_equal(np.asarray(acc2).view(np.uint32), np.asarray(acc).view(np.uint32)) and
                          n2 == spp and fp2 == fp and
                          np.array_equal(P.read_pfm(ppath).view(np.uint32), hdr.view(np.uint32)))
        ctx.free_scene(sid)
        log(f"io: {spp} spp in one call {one_ms:.1f} ms, in {spp} progressive calls {prog_ms:.1f} ms, identical "
            f"{same_img}; accumulation file save {acc_save_ms:.1f} ms, load {acc_load_ms:.1f} ms; PFM {pfm_ms:.1f} ms")

        # The drop-in CLI on the example scenes (a process each: scene load,
        # tree build, upload, render, BMP), against the golden images.
        gold = np.load(os.path.join(ROOT, "tests", "golden", "ref_images.npz"))
        sdir = os.path.join(ROOT, "tests", "golden", "scenes")
        cli = []
        for name in sorted(f for f in os.listdir(sdir) if f.endswith(".json")):
            out = os.path.join(tmp, name + ".bmp")
            t0 = time.perf_counter()
            r = subprocess.run([P.CLI, os.path.join(sdir, name), out], capture_output=True, text=True, timeout=300)
            wall = (time.perf_counter() - t0) * 1e3
            key = f"{name}:exact"
            same = bool(np.array_equal(bmp_rgb(out), gold[key])) if r.returncode == 0 and key in gold.files else None
            cli.append({"scene": name, "exit": r.returncode, "wall_ms": round(wall, 1), **cli_stage_ms(r.stdout),
                        "equals_golden": same})
            log(f"io: Pray {name}: exit {r.returncode}, {wall:.0f} ms {cli_stage_ms(r.stdout)}, equals golden {same}")
        cpu = None if args.no_cpu_baseline else cli_cpu_baseline(sdir, [c["scene"] for c in cli], gold, tmp)
        print(json.dumps({
            "metric": "flat tree file and progressive path tracing", "n_gpus": 1, "data": "synthetic",
            "flat_tree_file": {"tree": f"terrain {args.grid}x{args.grid}, binned SAH built on the CPU",
                               "build_s": round(build_s, 3), "save_s": round(save_s, 3), "bytes": size,
                               "load_s": round(load_s, 3), "upload_s": round(upload_s, 3),
                               "identical_traces": same_tree},
            "progressive": {"scene": f"Cornell room, {room.num_tris} triangles, {w}x{h}, depth {depth}",
                            "spp": spp, "one_call_ms": round(one_ms, 2), "progressive_ms": round(prog_ms, 2),
                            "per_call_overhead_ms": round((prog_ms - one_ms) / spp, 3),
                            "identical_image": same_img, "accum_save_ms": round(acc_save_ms, 2),
                            "accum_load_ms": round(acc_load_ms, 2), "pfm_write_ms": round(pfm_ms, 2),
                            "files_round_trip": same_files},
            "cli": cli,
            "cpu_baseline": cpu,
        }), flush=True)
    finally:
        ctx.close()
        shutil.rmtree(tmp, ignore_errors=True)


def build_workload(P, args):
    """The KD-tree builds on the GPU, timed on the terrain: the binned SAH
    build and the Morton hierarchy, each five times (the median of the last
    four; the first loads the kernels).  The CPU's binned SAH builder, which
    makes the same tree, is timed beside them, and closest-hit traces of
    2^20 rays over the GPU- and CPU-built SAH trees must agree bit for bit.
    One JSON line."""
    ctx = P.Context(0)
    configure(ctx, args)
    try:
        scene = P.Scene.terrain(args.grid, args.seed)
        abc, _, _ = scene.triangles()
        gpu, sids = {}, {}
        for name, method in (("binned_sah", P.BUILD_SAH), ("morton", P.BUILD_MORTON)):
            runs = []
            for i in range(5):
                t0 = time.perf_counter()
                sid, ms = ctx.build_scene(abc, method=method, leaf_size=args.leaf_size, max_depth=args.max_depth,
                                          traversal_cost=args.traversal_cost)
                ctx.sync()
                runs.append((ms, (time.perf_counter() - t0) * 1e3))
                if i < 4:
                    ctx.free_scene(sid)
            sids[name] = sid
            lay = ctx.scene_layout(sid)
            gpu[name] = {"build_ms": round(float(np.median([r[0] for r in runs[1:]])), 3),
                         "wall_ms_with_upload": round(float(np.median([r[1] for r in runs[1:]])), 3),
                         "first_build_ms": round(runs[0][0], 3), "depth": int(lay["max_depth"]),
                         "nodes": len(lay["nodes"]) // 2, "references": len(lay["refs"])}
            log(f"build: GPU {name} {gpu[name]}")
        del abc
        t0 = time.perf_counter()
        tree = scene.build(2, max_depth=args.max_depth, leaf_size=args.leaf_size, traversal_cost=args.traversal_cost)
        cpu_s = time.perf_counter() - t0
        sid_c = ctx.upload(tree, shading=False)
        rays = ctx.gen_rays(args.seed, P.scene_box_expanded(tree.box()), 1 << 20)
        ig, tg = ctx.trace(sids["binned_sah"], rays, P.TRACE_CLOSEST, args.det_eps)
        ic, tc = ctx.trace(sid_c, rays, P.TRACE_CLOSEST, args.det_eps)
        same = bool(np.array_equal(ig, ic) and np.array_equal(tg.view(np.uint32), tc.view(np.uint32)))
        log(f"build: CPU binned SAH {cpu_s:.2f} s, identical traces {same}")
        print(json.dumps({
            "metric": "KD-tree build time on the GPU", "n_gpus": 1, "data": "synthetic",
            "config": {"workload": f"terrain {args.grid}x{args.grid} cells = {scene.num_tris} triangles, depth cap "
                                   f"{args.max_depth}, leaf size {args.leaf_size}, traversal cost "
                                   f"{args.traversal_cost:g}, 64 bins per axis"},
            "gpu": gpu,
            "cpu_binned_sah": {"build_s": round(cpu_s, 3), "cores": int(os.environ.get("OMP_NUM_THREADS",
                                                                                         os.cpu_count() or 1))},
            "identical_traces": same,
        }), flush=True)
    finally:
        ctx.close()


def cli_stage_ms(text):
    """Stage times a CLI run prints (preprocessing, rendering, total), in ms."""
    import re

    out = {}
    for key, pat in (("preprocess_ms", r"Preprocessing time: (\d+) ms"), ("render_ms", r"Finished rendering\. \((\d+) ms\)"),
                     ("total_ms", r"Total time: (\d+) ms")):
        m = re.search(pat, text)
        out[key] = int(m.group(1)) if m else None
    return out


def cli_cpu_baseline(sdir, names, gold, tmp):
    """The reference's own CLI, compiled from its sources into the oracle's
    build folder, on the ray-traced example scenes (the path-traced one takes
    minutes on the CPU): wall time per scene and whether its image equals the
    golden one.  None when that binary was not built."""
    import oracle as O

    if not os.path.exists(O.REF_CLI):
        return None
    runs = []
    for name in names:
        if f"{name}:exact" not in gold.files:
            continue
        out = os.path.join(tmp, name + ".ref.bmp")
        t0 = time.perf_counter()
        r = subprocess.run([O.REF_CLI, os.path.join(sdir, name), out], capture_output=True, text=True, timeout=300)
        wall = (time.perf_counter() - t0) * 1e3
        same = bool(np.array_equal(bmp_rgb(out), gold[f"{name}:exact"])) if r.returncode == 0 else None
        runs.append({"scene": name, "exit": r.returncode, "wall_ms": round(wall, 1), **cli_stage_ms(r.stdout),
                     "equals_golden": same})
        log(f"io: reference Pray (CPU) {name}: exit {r.returncode}, {wall:.0f} ms {cli_stage_ms(r.stdout)}")
    return {"kind": "reference", "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "sample": "the reference CLI built here without FMA contraction, ray-traced example scenes, "
                      f"OpenMP on {cpu_model()}", "runs": runs}


def bmp_rgb(path):
    """A 24-bit BMP as (h, w, 3) RGB, top row first."""
    data = open(path, "rb").read()
    off = int.from_bytes(data[10:14], "little")
    w, h = (int.from_bytes(data[a:a + 4], "little", signed=True) for a in (18, 22))
    stride = (3 * w + 3) // 4 * 4
    rows = np.frombuffer(data, np.uint8, stride * abs(h), off).reshape(abs(h), stride)[:, :3 * w]
    rgb = rows.reshape(abs(h), w, 3)[:, :, [2, 1, 0]]
    return np.ascontiguousarray(rgb[::-1] if h > 0 else rgb)


def build_tree(P, ctx, scene, args, which):
    """Builds and uploads the terrain's tree: (scene id, tag, description,
    build info, (depth, nodes, references), raw box)."""
    params = (f"depth cap {args.max_depth}, leaf size {args.leaf_size}, traversal cost "
              f"{args.traversal_cost:g}")
    if which == "gpu-sah":
        abc, _, _ = scene.triangles()
        builds = []
        for _ in range(3):  # the median of three builds
            t1 = time.time()
            sid, bms = ctx.build_scene(abc, method=P.BUILD_SAH, leaf_size=args.leaf_size, max_depth=args.max_depth,
                                       traversal_cost=args.traversal_cost)
            builds.append((bms, time.time() - t1))
            if len(builds) < 3:
                ctx.free_scene(sid)
        del abc
        lay = ctx.scene_layout(sid)
        info = {"where": "GPU", "method": "binned SAH, 64 bins per axis, level by level",
                "build_ms": round(sorted(b[0] for b in builds)[1], 3),
                "build_wall_s_with_upload": round(sorted(b[1] for b in builds)[1], 3)}
        return (sid, f"gpusah-d{args.max_depth}-leaf{args.leaf_size}-ct{args.traversal_cost:g}",
                f"{TREES[which]}, {params}", info,
                (lay["max_depth"], len(lay["nodes"]) // 2, len(lay["refs"])), lay["box"])
    t1 = time.time()
    if which == "ref-sah":
        tree = scene.build(1)
        tag, desc, method = "sah-d20-leaf32-ct5", TREES[which], "exact-event SAH (the reference's sweep)"
    else:
        kind, tag, method = (2, "binsah", "binned SAH, 64 bins per axis") if which == "cpu-sah" else \
            (1, "sah", "exact-event SAH (the reference's sweep)")
        tree = scene.build(kind, max_depth=args.max_depth, leaf_size=args.leaf_size,
                           traversal_cost=args.traversal_cost)
        tag = f"{tag}-d{args.max_depth}-leaf{args.leaf_size}-ct{args.traversal_cost:g}"
        desc = f"{TREES[which]}, {params}"
    build_s = time.time() - t1
    st = tree.stats()
    sid = ctx.upload(tree, shading=False)
    return (sid, tag, desc, {"where": "CPU", "method": method, "build_s": round(build_s, 2)},
            (st["maxdepth"], st["flat_nodes"], st["flat_refs"]), tree.box())


def gpu_sah_scene(P, ctx, scene, args):
    """Uploads a scene with its SAH KD tree built on the GPU (with shading
    data): (scene id, device build milliseconds)."""
    abc, col, emit = scene.triangles()
    return ctx.build_scene(abc, col, emit, method=P.BUILD_SAH, leaf_size=args.leaf_size, max_depth=args.max_depth,
                           traversal_cost=args.traversal_cost)


def configure(ctx, args):
    """Applies --no-graphs and --path-budget to a new context."""
    ctx.use_graphs(not args.no_graphs)
    if args.path_budget >= 0:
        ctx.set_path_budget(args.path_budget)


# Tuning fields that select the traversal kernel's instance (the others only
# change its launch or refill threshold).
KERNEL_FIELDS = ("layout", "stack_lds", "leaf_ilp", "node_levels")


def pmc_traffic(args, tree_tag, tun):
    """HBM bytes per launch of the traversal kernel instance `tun` selects,
    from profiles/trace_closest_terrain<grid>_<tree>_pmc.json (one entry per
    profiled tuning), or None when that instance was not profiled."""
    path = os.path.join(ROOT, "profiles", f"trace_closest_terrain{args.grid}_{tree_tag}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        entries = json.load(fh).get("entries", [])
    for e in entries:
        et = e.get("tuning", {})
        if all(et.get(k) == tun[k] for k in KERNEL_FIELDS) and (et.get("leaf_batch", 0) > 0) == (tun["leaf_batch"] > 0):
            return e.get("hbm_bytes_per_launch")
    return None


def kd_bytes(cnt):
    """Algorithmic bytes per ray of a KD traversal with (nodes, references)
    per ray `cnt`: ray 24 + result 8 + 8 per node + 40 per reference."""
    return 24 + 8 + 8 * float(cnt[0]) + 40 * float(cnt[1])


def boxed_bytes(cnt):
    """Bytes per ray of a boxed traversal with (children tested, references)
    per ray `cnt`: ray 24 + result 8 + 28 per child (its box and word) + 40
    per reference."""
    return 24 + 8 + 28 * float(cnt[0]) + 40 * float(cnt[1])


def parse_tuning(P, text):
    """--tuning's comma-separated values as set_tuning keywords."""
    return dict(zip(P.TUNING_FIELDS, [int(x) for x in text.split(",")]))


def P_camera(P, scene, side):
    """The scene's camera at a square image of `side` pixels."""
    inf = scene.info()
    scene.set_render(side, side, inf["spp"], inf["max_depth"], inf["rr_depth"])
    return scene.camera()


def n_rays_for_tuning(args):
    """The autotuner times the workload's own stream size (capped at 2^24)."""
    return min(args.rays, 1 << 24)


def equal_rmse_time(points, target):
    """Time at which a method whose runs gave (ms, rmse) points reaches
    `target`: log-log interpolation between the two nearest points (RMSE of
    an unbiased estimator falls as time^-1/2), extrapolated at the ends."""
    pts = sorted(points)
    lt = np.log([p[0] for p in pts])
    lr = np.log([p[1] for p in pts])
    k = int(np.searchsorted(-lr, -np.log(target)))
    k = min(max(k, 1), len(pts) - 1)
    slope = (lt[k] - lt[k - 1]) / (lr[k] - lr[k - 1])
    return float(np.exp(lt[k - 1] + slope * (np.log(target) - lr[k - 1])))


def rt_workload(P, args):
    """--workload rt: ray_tracer_leg as one JSON line of its own."""
    ctx = P.Context(0)
    configure(ctx, args)
    r = ray_tracer_leg(P, ctx, args)
    ctx.close()
    print(json.dumps({"metric": "Mrays/s ray tracing (primary + shadow rays), spheres scene",
                      "value": r.pop("Mrays_per_s"), "unit": "Mrays/s", "n_gpus": 1, "higher_is_better": True,
                      "dtype": "f32", "data": "synthetic", "config": {"workload": r.pop("workload")}, **r}),
          flush=True)


def ray_tracer_leg(P, ctx, args):
    """The ray tracer (pray_render_rt) on the spheres scene: a level-4
    icosphere, 16 level-3 icospheres and a ground quad (25,602 triangles),
    median KD tree, 512x512, fov 60, one point light.  --steps calls are
    timed after --warmup (each call returns the 8-bit image to the host);
    rays are the primary rays plus one shadow ray per lit hit.  The scene is
    also written as JSON + OBJ and rendered by the drop-in CLI and, as the
    CPU baseline, by the reference's own CLI compiled here; their images are
    compared byte for byte."""
    import shutil
    import tempfile

    scene = P.Scene.spheres(1)
    tree = scene.build(0)
    sid = ctx.upload(tree)
    if args.rt_tuning:
        ctx.set_tuning(sid, **parse_tuning(P, args.rt_tuning))
    tun = ctx.get_tuning(sid)
    inf = scene.info()
    w, h = inf["width"], inf["height"]
    cam, lights, bg, eps = scene.camera(), scene.lights(), inf["background"], float(inf["det_eps"])
    # the frame loop fills the same page-locked host image and hit buffers
    # every frame
    frame_out = (ctx.host_array((h, w, 3), np.uint8), ctx.host_array(w * h, np.int32))
    for _ in range(max(1, args.warmup)):
        rgb, hits = ctx.render_rt(sid, cam, w, h, lights, bg, eps, want_hits=True, out=frame_out)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        rgb, hits = ctx.render_rt(sid, cam, w, h, lights, bg, eps, want_hits=True, out=frame_out)
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    rays = w * h + int((hits >= 0).sum()) * len(lights)
    # The frame's closest-hit launch alone, on the same camera rays (the
    # frame builds them on the device, bit-identical): work counts, then K
    # launches timed with HIP events on the kernel's stream.
    npix = w * h
    d_rays = ctx.to_device(scene.prim_rays())
    d_idx, d_t = ctx.alloc(npix, np.int32), ctx.alloc(npix, np.float32)
    d_cnt = ctx.alloc((npix, 2), np.uint32)
    ctx.trace_counts_device(sid, d_rays.ptr, npix, d_idx.ptr, d_t.ptr, d_cnt.ptr, P.TRACE_CLOSEST, eps)
    ctx.sync()
    cnt = d_cnt.numpy().astype(np.int64).sum(0)
    assert np.array_equal(d_idx.numpy(), hits)
    nodes_per_ray, refs_per_ray = cnt[0] / npix, cnt[1] / npix
    # the median tree's own plane-form traversal tests ~150 nodes and ~740
    # triangles per ray; the boxed forms (its default) are priced by their
    # own records: 28 bytes per child tested
    boxed = tun["node_levels"] in (3, 4)
    bytes_per_ray = (boxed_bytes if boxed else kd_bytes)(cnt / npix)
    for _ in range(max(1, args.warmup)):
        ctx.trace_device(sid, d_rays.ptr, npix, d_idx.ptr, d_t.ptr, P.TRACE_CLOSEST, eps)
    ctx.sync()
    ctx.record(0)
    for _ in range(args.steps):
        ctx.trace_device(sid, d_rays.ptr, npix, d_idx.ptr, d_t.ptr, P.TRACE_CLOSEST, eps)
    ctx.record(1)
    kernel_ms = ctx.elapsed_ms(0, 1) / args.steps
    ctx.sync()
    for b in (d_rays, d_idx, d_t, d_cnt):
        b.free()
    achieved = bytes_per_ray * npix / (kernel_ms * 1e-3) / 1e9
    ctx.free_scene(sid)
    log(f"rt: {scene.num_tris} triangles {w}x{h}: {ms:.3f} ms per frame (median of {args.steps}), "
        f"{rays} rays, {rays / ms / 1e3:.1f} Mrays/s; closest-hit launch {kernel_ms:.4f} ms, "
        f"nodes/ray {nodes_per_ray:.1f} refs/ray {refs_per_ray:.1f}")
    tmp = tempfile.mkdtemp(prefix="pray_rt_")
    try:
        path = os.path.join(tmp, "spheres.json")
        scene.write_json(path)
        out = os.path.join(tmp, "gpu.bmp")
        t0 = time.perf_counter()
        r = subprocess.run([P.CLI, path, out], capture_output=True, text=True, timeout=300)
        cli = {"exit": r.returncode, "wall_ms": round((time.perf_counter() - t0) * 1e3, 1), **cli_stage_ms(r.stdout)}
        log(f"rt: drop-in CLI {cli}")
        cpu = None
        if not args.no_cpu_baseline:
            cpu = rt_cpu_baseline(path, out if r.returncode == 0 else None, tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {
        "Mrays_per_s": round(rays / ms / 1e3, 2),
        "workload": f"ray tracer, {scene.num_tris} triangles (1 level-4 and 16 level-3 icospheres, a ground quad), "
                    f"median KD tree, {w}x{h}, fov {inf['fov']}, {len(lights)} light",
        "ms_per_frame": round(ms, 4), "rays_per_frame": rays, "tuning": tun,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": {4: "k_trace (closest hit, 4-wide boxed records)",
                                3: "k_trace (closest hit, boxed children pairs)",
                                2: "k_trace_q (closest hit, two tree levels per node step)"}.get(
                                    tun["node_levels"], "k_trace (closest hit)"),
                     "kernel_ms": round(kernel_ms, 4), "bytes_per_ray": round(float(bytes_per_ray), 1),
                     "bytes_basis": ("24 + 8 + 28 per boxed child tested + 40 per reference" if boxed else
                                     "24 + 8 + 8 per node + 40 per reference"),
                     ("children_per_ray" if boxed else "nodes_per_ray"): round(float(nodes_per_ray), 2),
                     "refs_per_ray": round(float(refs_per_ray), 2),
                     "note": "the scene's records (1.3 MB) stay in L2; the launch of 262,144 rays is bound by "
                             "its longest rays' chains of dependent loads, not by HBM"},
        "cli": cli, "cpu_baseline": cpu}


def rt_cpu_baseline(scene_json, gpu_bmp, tmp):
    """The reference's own CLI, compiled from its sources into the oracle's
    build folder, on the same scene file: its times and whether its image
    equals the drop-in CLI's byte for byte.  None when that binary was not
    built."""
    import oracle as O

    if not os.path.exists(O.REF_CLI):
        return None
    out = os.path.join(tmp, "cpu.bmp")
    t0 = time.perf_counter()
    r = subprocess.run([O.REF_CLI, scene_json, out], capture_output=True, text=True, timeout=600)
    wall = (time.perf_counter() - t0) * 1e3
    same = None
    if r.returncode == 0 and gpu_bmp:
        same = open(out, "rb").read() == open(gpu_bmp, "rb").read()
    st = cli_stage_ms(r.stdout)
    log(f"rt: reference CLI (CPU) exit {r.returncode}, {wall:.0f} ms {st}, image equal to the drop-in CLI's: {same}")
    return {"kind": "reference", "cores": USABLE_CPUS, "exit": r.returncode, "wall_ms": round(wall, 1), **st,
            "equal_to_gpu_cli_bmp": same,
            "sample": f"the reference CLI built here without FMA contraction on the same scene file (it ignores "
                      f"the file's tree key and builds its SAH tree, as its policy gives for 20k-1M triangles), "
                      f"OpenMP on {cpu_model()}"}


def pt_workload(P, args):
    """The Cornell room (100k triangles, depth 8, roulette from bounce 3) at
    --cache-size^2: one warm-up and one timed plain render of --pt-spp
    samples per pixel, then one radiance-cache render of --pt-frames frames.
    One JSON line with the times and rays; meant for profilers, which see
    the kernels of each pass (with --no-graphs, launched one by one)."""
    ctx = P.Context(0)
    configure(ctx, args)
    w = h = args.cache_size
    scene = P.Scene.cornell(1, 100_000)
    tree = scene.build(1)  # the reference's SAH KD tree (its event sweep and parameters)
    sid = ctx.upload(tree)
    cam, bg = scene.camera(), scene.info()["background"]
    depth, rr = 8, 3
    tun = ctx.get_tuning(sid)
    if args.tuning:
        ctx.set_tuning(sid, **parse_tuning(P, args.tuning))
        tun = ctx.get_tuning(sid)
    # the warm-up at the timed sample count grows the render workspace to
    # its size; the median of three timed renders
    ctx.render_pt(sid, cam, w, h, args.pt_spp, depth, rr, 7, bg, 1e-9, want_rgb=False)
    ctx.sync()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        img, _ = ctx.render_pt(sid, cam, w, h, args.pt_spp, depth, rr, 7, bg, 1e-9, want_rgb=False)
        times.append((time.perf_counter() - t0) * 1e3)
    plain_ms = float(np.median(times))
    rays = ctx.last_render_rays()
    segs = rays + (args.pt_spp - 1) * w * h  # every sample's camera ray counted
    log(f"pt: plain {args.pt_spp} spp {plain_ms:.1f} ms, {rays} rays traced ({rays / plain_ms / 1e3:.1f} Mrays/s), "
        f"{segs} path segments ({segs / plain_ms / 1e3:.1f} M/s)")
    c = P.Cache(ctx, seed=3, learning_rate=1e-2, max_batch=1 << 16)
    t0 = time.perf_counter()
    _, _, st = c.render_pt(sid, cam, w, h, args.pt_frames, depth, rr, 7, bg, 1e-9, train_paths=1 << 16,
                           cache_bounce=1, min_steps=4, samples_per_frame=args.cache_spf)
    cache_ms = (time.perf_counter() - t0) * 1e3
    c.close()
    log(f"pt: cache {args.pt_frames} frames of {args.cache_spf} samples {cache_ms:.1f} ms, "
        f"{int(st[:, 1].sum())} with the cache")
    ctx.free_scene(sid)
    ctx.close()
    print(json.dumps({"metric": "path tracing, Cornell room", "n_gpus": 1, "data": "synthetic",
                      "config": {"workload": f"Cornell room, {scene.num_tris} triangles, {w}x{h}, depth {depth}, "
                                             f"roulette from bounce {rr}, the reference's SAH KD tree"},
                      "graphs": not args.no_graphs, "path_budget": args.path_budget, "tuning": tun,
                      "plain": {"spp": args.pt_spp, "ms": round(plain_ms, 2), "rays": rays,
                                "Mrays_per_s": round(rays / plain_ms / 1e3, 2), "path_segments": segs,
                                "Msegments_per_s": round(segs / plain_ms / 1e3, 2),
                                "checksum": float(img.astype(np.float64).sum())},
                      "cache": {"frames": args.pt_frames, "samples_per_frame": args.cache_spf,
                                "ms": round(cache_ms, 2),
                                "frames_with_cache": int(st[:, 1].sum())}}), flush=True)


def radiance_cache_leg(P, ctx, args):
    """Equal-RMSE wall time of the cache path tracer against the plain
    wavefront path tracer on the Cornell scene (100k triangles, depth 8,
    roulette from bounce 3), both measured against a plain reference with
    --ref-spp samples per pixel, RMSE on linear HDR."""
    w = h = args.cache_size
    scene = P.Scene.cornell(1, 100_000)
    tree = scene.build(1)  # the reference's SAH KD tree (its event sweep and parameters)
    sid = ctx.upload(tree)
    inf = scene.info()
    cam, bg = scene.camera(), inf["background"]
    depth, rr = 8, 3
    tun = None
    if not args.no_autotune:  # tuned on this path-tracing workload, for both tracers
        tun = ctx.autotune_pt(sid, cam, w, h, depth, rr, 7, bg, 1e-9)
        log(f"cache leg: tuning {tun}")
    t0 = time.perf_counter()
    ref, _ = ctx.render_pt(sid, cam, w, h, args.ref_spp, depth, rr, 1000003, bg, 1e-9, want_rgb=False)
    log(f"cache leg: reference {args.ref_spp} spp in {time.perf_counter() - t0:.1f}s")

    def rmse(img):
        return float(np.sqrt(np.mean((img.astype(np.float64) - ref) ** 2)))

    plain = []
    for spp in (16, 32, 64, 128, 256, 512, 1024, 2048):
        t0 = time.perf_counter()
        img, _ = ctx.render_pt(sid, cam, w, h, spp, depth, rr, 7, bg, 1e-9, want_rgb=False)
        ms = (time.perf_counter() - t0) * 1e3
        plain.append({"spp": spp, "ms": round(ms, 2), "rmse": round(rmse(img), 6)})
        log(f"cache leg: plain {spp} spp {ms:.1f} ms rmse {plain[-1]['rmse']:.5f}")
    cache = []
    # plain frames before the cache takes over; training samples per frame (2^16, the workload's);
    # samples per pixel in a frame (one training step each)
    warm, train, spf = 4, 1 << 16, args.cache_spf
    for total in (32, 64, 128, 256, 512, 1024):
        frames = max(1, total // spf)
        c = P.Cache(ctx, seed=3, learning_rate=1e-2, max_batch=1 << 16)
        t0 = time.perf_counter()
        img, _, st = c.render_pt(sid, cam, w, h, frames, depth, rr, 7, bg, 1e-9, train_paths=train,
                                 cache_bounce=1, min_steps=warm, samples_per_frame=spf)
        ms = (time.perf_counter() - t0) * 1e3
        c.close()
        cache.append({"samples": frames * spf, "frames": frames, "ms": round(ms, 2), "rmse": round(rmse(img), 6),
                      "frames_with_cache": int(st[:, 1].sum()), "final_loss": round(float(st[-1, 0]), 4)})
        log(f"cache leg: cache {frames} frames x {spf} samples {ms:.1f} ms rmse {cache[-1]['rmse']:.5f}")
    pp = [(p["ms"], p["rmse"]) for p in plain]
    for cpt in cache:
        cpt["plain_ms_at_equal_rmse"] = round(equal_rmse_time(pp, cpt["rmse"]), 2)
        cpt["speedup"] = round(cpt["plain_ms_at_equal_rmse"] / cpt["ms"], 3)
    ctx.free_scene(sid)
    best = cache[-1]
    return {
        "scene": f"Cornell room, {scene.num_tris} triangles, {w}x{h}, depth {depth}, roulette from bounce {rr}, "
                 f"the reference's SAH KD tree built on the CPU (its event sweep and parameters)",
        "reference_spp": args.ref_spp, "tuning": tun,
        "cache": {"network": "192 (172 used) -> 5 x 64 ReLU -> 3, BF16 matrix cores, FP32 Adam lr 1e-2",
                  "training_samples_per_frame": train, "samples_per_frame": spf, "terminates_at_hit": 2,
                  "warmup_frames": warm},
        "plain_runs": plain,
        "cache_runs": cache,
        "equal_rmse_speedup": best["speedup"],
        "at_rmse": best["rmse"],
    }


def cpu_baseline(scene, args, box, gpu_idx):
    """The reference CPU path restated (oracle): OpenMP schedule(dynamic,10)
    over 8-ray AVX packets, exhaustive recursive traversal of the median KD tree
    (the reference's choice at >= 1M triangles), AVX Moeller-Trumbore.  Brute
    force over every triangle (bf_search.cpp:26-40) is timed beside it on the
    first 4096 rays, and the reference's own compiled traversal on 2^20."""
    import oracle as O

    cpu = cpu_info()
    threads = cpu["usable_cpus"]  # one OpenMP thread per CPU this process may run on
    tree = scene.build(0)
    ex = tree.export()
    fl = tree.flat()
    n = min(args.cpu_rays, args.rays)
    rays = O.gen_rays(args.seed, box, n)
    with omp_master_pinned():
        idx, t, sec = O.ref_closest_avx(ex, fl["a"], fl["ab"], fl["ac"], rays, args.det_eps, threads)
    log(f"cpu baseline: {n} rays in {sec:.2f}s on {threads} threads ({cpu['model']}, "
        f"{cpu.get('physical_cores', '?')} physical cores, {cpu.get('logical_cpus', '?')} logical CPUs, "
        f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} OMP_PLACES={os.environ.get('OMP_PLACES')})")
    nb = min(4096, n)
    t0 = time.perf_counter()
    bidx, _ = O.bf_closest(fl["a"], fl["ab"], fl["ac"], rays[:nb], args.det_eps)
    bsec = time.perf_counter() - t0
    same = int((bidx == gpu_idx[:nb]).sum())
    log(f"brute force: {nb} rays in {bsec:.2f}s, {same}/{nb} hit indices equal to the GPU's")
    # The reference's own code (compiled from its sources into the oracle's
    # build folder): its median KD tree and AVX packet traversal.  Its
    # determinant cull is fixed at 1e-4, which no terrain triangle passes, so
    # it compares with the GPU's reference_det_eps run, not with `value`.
    ref_code = None
    if os.path.exists(O.REF_LIB):
        abc, _, _ = scene.triangles()
        with omp_master_pinned():
            t0 = time.perf_counter()
            db = O.RefDB(abc, 0)
            rbuild = time.perf_counter() - t0
            del abc
            nr = min(1 << 20, n)
            ridx, _, rsec = db.intersect_avx(rays[:nr], threads)
        db.close()
        ref_code = {"sample": f"first {nr} rays", "det_eps": 1e-4, "hits": int((ridx >= 0).sum()),
                    "Mrays_per_s": round(nr / rsec / 1e6, 4), "build_s": round(rbuild, 2), "cores": threads,
                    "what": "the reference's KDTree (median) and AVX packet intersect compiled here from its "
                            "sources; its det cull is fixed at 1e-4"}
        log(f"reference code: {nr} rays in {rsec:.2f}s ({ref_code['hits']} hits), median tree built in "
            f"{rbuild:.1f}s")
    return {
        "value": round(n / sec / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} rays of the same stream, same det_eps, restated reference AVX packet traversal over "
                  f"the reference's median KD tree (identical to the reference build), {threads} OpenMP threads on "
                  f"{cpu['model']}",
        "cpu": {**cpu, "threads_used": threads, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
                "omp_places": os.environ.get("OMP_PLACES")},
        "reference_code": ref_code,
        "brute_force": {"rays": nb, "seconds": round(bsec, 2), "Mrays_per_s": round(nb / bsec / 1e6, 6),
                        "same_hit_index_as_gpu": same,
                        "what": "scalar Moeller-Trumbore over every triangle, lowest index wins ties, "
                                f"{threads} OpenMP threads"},
    }


if __name__ == "__main__":
    main()
