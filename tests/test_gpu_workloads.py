"""Parity at the workloads' own sizes and settings, through the C ABI, against
the oracle:
  - closest and any hit of the 2^24-ray incoherent stream over the
    2048x2048-cell terrain (8,388,608 triangles) with the benchmark's tree
    (the reference's SAH KD tree built on the CPU: its event sweep, depth cap
    20, leaf 32, costs 5 / 1.5) and its autotuned setting (4-wide boxed node
    steps): a prefix and a stride of the stream against the
    CPU twin of the traversal, 4096 rays against brute force
    (bf_search.cpp:26-40), and the whole stream equal across settings;
  - the 100,248-triangle Cornell room path traced at depth 8 with roulette
    from bounce 3 (images and per-bounce queue sizes);
  - one batch spanning the multi-scene workload's size buckets (1k to >1M
    triangles, 128^2 to 1024^2 images, trees built on the GPU, depth 6), every
    scene's image against the restated path tracer."""
import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O
from conftest import first_diff, say

pytestmark = pytest.mark.gpu

# The setting the autotuner picks for the benchmark's stream (4-wide boxed
# node steps, ring of 8, ILP 2, refill at 24, leaf steps once 16 lanes
# wait), compared with the oracle first; the others must give the same hits.
BENCH_TUNING = dict(layout=P.LAYOUT_PACKED, stack_lds=8, leaf_ilp=2, fetch_min=24, occupancy=100, wg_waves=1,
                    leaf_batch=16, node_levels=4)


def test_terrain2048_full_stream_equals_oracle(ctx):
    say("terrain 2048: scene and the reference's SAH tree")
    scene = P.Scene.terrain(2048, 2024)
    tree = scene.build(1)
    flat = tree.flat()
    assert flat["a"].shape[0] == 8_388_608
    sid = ctx.upload(tree, shading=False)
    n = 1 << 24
    box = P.scene_box_expanded(tree.box())
    bufs = []
    try:
        d_rays = ctx.alloc((n, 6), np.float32)
        d_idx, d_t = ctx.alloc(n, np.int32), ctx.alloc(n, np.float32)
        d_lim = ctx.to_device(np.full(n, np.finfo(np.float32).max, np.float32))
        bufs = [d_rays, d_idx, d_t, d_lim]
        ctx.gen_rays_device(2024, box, n, d_rays.ptr)
        rays = d_rays.numpy()
        # the device stream is the oracle's stream
        assert np.array_equal(rays[:1 << 16].view(np.uint32), O.gen_rays(2024, box, 1 << 16).view(np.uint32))
        sample = np.concatenate([np.arange(1 << 16), np.arange(1 << 16, n, 251)])
        first = None
        for tun in (BENCH_TUNING, dict(BENCH_TUNING, stack_lds=8, leaf_ilp=1, fetch_min=24, node_levels=1),
                    dict(BENCH_TUNING, stack_lds=8, leaf_batch=16, node_levels=1),
                    dict(layout=P.LAYOUT_INDEXED, stack_lds=8, leaf_ilp=2, fetch_min=32, wg_waves=2, leaf_batch=0,
                         node_levels=1),
                    dict(BENCH_TUNING, stack_lds=6, leaf_batch=0, node_levels=2), dict(BENCH_TUNING, leaf_ilp=1),
                    dict(BENCH_TUNING, leaf_batch=0, node_levels=3)):
            ctx.set_tuning(sid, **tun)
            say(f"terrain 2048: closest hit, 2^24 rays, {tun}")
            ctx.trace_device(sid, d_rays.ptr, n, d_idx.ptr, d_t.ptr, P.TRACE_CLOSEST, 1e-12)
            ctx.sync()
            gi, gt = d_idx.numpy(), d_t.numpy()
            if first is None:
                first = (gi, gt)
                oi, ot = O.flat_trace(flat, rays[sample], 0, 1e-12)
                assert first_diff(gi[sample], oi) is None, first_diff(gi[sample], oi)
                assert np.array_equal(gt[sample].view(np.uint32), ot.view(np.uint32))
                assert 0.1 < (gi >= 0).mean() < 0.5
            else:
                assert first_diff(gi, first[0]) is None, (tun, first_diff(gi, first[0]))
                assert np.array_equal(gt.view(np.uint32), first[1].view(np.uint32)), tun
        say("terrain 2048: brute force over the first 4096 rays")
        bi, bt = O.bf_closest(flat["a"], flat["ab"], flat["ac"], rays[:4096], 1e-12)
        assert first_diff(first[0][:4096], bi) is None, first_diff(first[0][:4096], bi)
        h = bi >= 0
        assert np.array_equal(first[1][:4096][h], bt[h])
        # any hit with unlimited distance at the benchmark's setting
        ctx.set_tuning(sid, **BENCH_TUNING)
        say("terrain 2048: any hit, 2^24 rays")
        ctx.trace_device(sid, d_rays.ptr, n, d_idx.ptr, d_t.ptr, P.TRACE_ANY, 1e-12, d_limit2=d_lim.ptr)
        ctx.sync()
        ai, at = d_idx.numpy(), d_t.numpy()
        assert np.array_equal(ai >= 0, first[0] >= 0)
        lim = np.full(len(sample), np.finfo(np.float32).max, np.float32)
        oa, _ = O.flat_trace(flat, rays[sample], 1, 1e-12, limit2=lim)
        assert np.array_equal(ai[sample] >= 0, oa >= 0)
        # the boxed node steps meet the occluders in another order than the
        # KD twin: every reported one is a genuine hit at the reported t
        for k in np.flatnonzero(ai[:4096] >= 0)[:512]:
            hit, tt = O.mt(rays[k], flat["a"][ai[k]], flat["ab"][ai[k]], flat["ac"][ai[k]], 1e-12)
            assert hit and tt > 0 and tt == at[k], k
    finally:
        for b in bufs:
            b.free()
        ctx.free_scene(sid)


@pytest.mark.parametrize("tree_kind", ["binned-sah-d40-leaf8", "reference-sah"])
def test_cornell100k_depth8_equals_oracle(ctx, tree_kind):
    s = P.Scene.cornell(1, 100_000)
    assert s.num_tris > 100_000
    t = s.build(2, max_depth=40, leaf_size=8, traversal_cost=2.0) if tree_kind == "binned-sah-d40-leaf8" \
        else s.build(1)
    f = t.flat()
    inf = s.info()
    w, h, spp, depth, rr, seed = 160, 120, 4, 8, 3, 7
    s.set_render(w, h, spp, depth, rr)
    cam = s.camera()
    bg = inf["background"]
    sid = ctx.upload(t)
    try:
        res = []
        for tun in (dict(layout=P.LAYOUT_PACKED, stack_lds=4, leaf_ilp=2, fetch_min=16, wg_waves=1, leaf_batch=0,
                         node_levels=1),
                    dict(layout=P.LAYOUT_INDEXED, stack_lds=8, leaf_ilp=1, fetch_min=32, wg_waves=2, leaf_batch=0,
                         node_levels=1),
                    dict(layout=P.LAYOUT_PACKED, stack_lds=6, leaf_ilp=2, fetch_min=16, wg_waves=1, leaf_batch=16,
                         node_levels=1),
                    dict(layout=P.LAYOUT_PACKED, stack_lds=6, leaf_ilp=2, fetch_min=16, wg_waves=1, leaf_batch=0,
                         node_levels=2),
                    dict(layout=P.LAYOUT_PACKED, stack_lds=6, leaf_ilp=2, fetch_min=16, wg_waves=1, leaf_batch=0,
                         node_levels=4)):
            ctx.set_tuning(sid, **tun)
            say(f"cornell 100k {tree_kind} {w}x{h} {spp} spp depth {depth} rr {rr} {tun}")
            hdr, rgb = ctx.render_pt(sid, cam, w, h, spp, depth, rr, seed, bg, 1e-9)
            res.append((hdr, rgb, ctx.last_render_queues()))
        say("cornell 100k: cache renderer with the cache held off equals the plain tracer")
        c = P.Cache(ctx, seed=3, max_batch=1 << 12)
        try:
            ch, _, st = c.render_pt(sid, cam, w, h, spp, depth, rr, seed, bg, 1e-9, train_paths=1 << 12,
                                    cache_bounce=1, max_loss=0.0, min_steps=1)
        finally:
            c.close()
    finally:
        ctx.free_scene(sid)
    oh, orgb, oq = O.render_pt(f, cam, w, h, spp, depth, rr, seed, bg, 1e-9, queues=True)
    for hdr, rgb, q in res:
        assert first_diff(hdr.view(np.uint32).ravel(), oh.view(np.uint32).ravel()) is None, \
            first_diff(hdr.ravel(), oh.ravel())
        assert np.array_equal(rgb, orgb)
        assert np.array_equal(q, oq), (q, oq)
    assert oq[0] == w * h * spp and oq[rr + 1] < oq[rr] and oq[-1] > 0
    assert np.array_equal(ch.view(np.uint32), oh.view(np.uint32)) and not st[:, 1].any()
    assert 0.05 < oh.mean() < 10


def test_cornell100k_full_size_split_invariance(ctx):
    """The Cornell room at the workload's own image size (1024^2, depth 8,
    roulette from bounce 3): samples traced all in one pass (6 x 1024^2 paths
    in flight), in passes of 2 (graph replays) and one by one give the same
    image bit for bit, the same traced rays and per-bounce queues, whose
    first entry is every path of every sample."""
    s = P.Scene.cornell(1, 100_000)
    t = s.build(2, max_depth=40, leaf_size=8, traversal_cost=2.0)
    inf = s.info()
    w = h = 1024
    spp, depth, rr, seed = 6, 8, 3, 7
    s.set_render(w, h, spp, depth, rr)
    cam = s.camera()
    sid = ctx.upload(t)
    out = []
    try:
        for budget in (1 << 24, 2 * w * h, 0):
            ctx.set_path_budget(budget)
            say(f"cornell 100k {w}x{h} {spp} spp, path budget {budget}")
            hdr, _ = ctx.render_pt(sid, cam, w, h, spp, depth, rr, seed, inf["background"], 1e-9, want_rgb=False)
            out.append((hdr, ctx.last_render_rays(), ctx.last_render_queues()))
    finally:
        ctx.set_path_budget(1 << 24)
        ctx.free_scene(sid)
    hdr0, rays0, q0 = out[0]
    # rays traced: the paths of every bounce after the first, and the camera
    # rays once (every sample of a pixel starts with its camera ray)
    assert q0[0] == w * h * spp and rays0 == int(q0.sum()) - int(q0[0]) + w * h and hdr0.mean() > 0.05
    for hdr, rays, q in out[1:]:
        assert np.array_equal(hdr.view(np.uint32), hdr0.view(np.uint32))
        assert rays == rays0 and np.array_equal(q, q0)


def gpu_flat(ctx, sid, scene):
    """The oracle's view of a scene whose tree was built on the GPU: the
    device layout plus the host's per-triangle data."""
    lay = ctx.scene_layout(sid)
    abc, col, emit = scene.triangles()
    ab, ac, nn = scene.derived()
    return {"nodes": lay["nodes"], "refs": lay["refs"], "a": np.ascontiguousarray(abc[:, :3]), "ab": ab, "ac": ac,
            "normal": nn, "color": col, "emit": emit, "box": lay["box"], "max_depth": lay["max_depth"]}


def test_batch_size_buckets_equal_oracle(ctx):
    """One batch over the multi-scene workload's buckets: trees built on the
    GPU as the benchmark builds them, 2 spp, depth 6, roulette from bounce 3."""
    spec = [(301, 1_000, 128), (302, 20_000, 256), (303, 250_000, 512), (304, 1_300_000, 1024), (305, 6_000, 256)]
    scenes, sids, cams = [], [], []
    try:
        for seed, tris, side in spec:
            s = P.Scene.random(seed, tris)
            abc, col, emit = s.triangles()
            say(f"batch: GPU SAH build of scene {seed} ({s.num_tris} triangles)")
            sid, _ = ctx.build_scene(abc, col, emit, method=P.BUILD_SAH, leaf_size=8, max_depth=40,
                                     traversal_cost=2.0)
            inf = s.info()
            s.set_render(side, side, inf["spp"], inf["max_depth"], inf["rr_depth"])
            scenes.append(s)
            sids.append(sid)
            cams.append(s.camera())
        assert max(s.num_tris for s in scenes) > 1_000_000
        sizes = [(side, side) for _, _, side in spec]
        seeds = [17 + k for k in range(len(spec))]
        bgs = [s.info()["background"] for s in scenes]
        spp, depth, rr = 2, 6, 3
        runs = []
        # The batch takes the setting of scene 3 (the largest image of the
        # bucket holding most pixels).
        for tun in (dict(layout=P.LAYOUT_PACKED, stack_lds=8, leaf_ilp=2, fetch_min=24, wg_waves=1, leaf_batch=0),
                    dict(layout=P.LAYOUT_PACKED, stack_lds=6, leaf_ilp=1, fetch_min=16, wg_waves=2, leaf_batch=16)):
            ctx.set_tuning(sids[3], **tun)
            say(f"batch of {len(spec)} scenes, {spp} spp, depth {depth}, {tun}")
            imgs, ms = ctx.render_pt_batch(sids, cams, sizes, spp, depth, rr, seeds, bgs, 1e-9)
            runs.append((imgs, ctx.last_render_rays()))
        total = 0
        for k, s in enumerate(scenes):
            say(f"batch: oracle render of scene {k}")
            f = gpu_flat(ctx, sids[k], s)
            oh, _, oq = O.render_pt(f, cams[k], sizes[k][0], sizes[k][1], spp, depth, rr, seeds[k], bgs[k], 1e-9,
                                    queues=True)
            total += int(oq.sum()) - int(oq[0]) + sizes[k][0] * sizes[k][1]  # camera rays traced once
            assert oh.mean() > 0
            for imgs, _ in runs:
                assert first_diff(imgs[k].view(np.uint32).ravel(), oh.view(np.uint32).ravel()) is None, \
                    (k, first_diff(imgs[k].ravel(), oh.ravel()))
        for _, rays in runs:
            assert rays == total
    finally:
        for sid in sids:
            ctx.free_scene(sid)


def test_batch_argument_limits(ctx):
    """The batch refuses what the single-scene renderer refuses (depth above
    64, more than 2^24 samples) before any device work."""
    s = P.Scene.random(41, 2000)
    t = s.build(1)
    sid = ctx.upload(t)
    try:
        cam = O.camera([0, 4, 9], [0, -0.45, -1], 60, 16, 16)
        bg = s.info()["background"]
        for spp, depth in ((1, 65), ((1 << 24) + 1, 2)):
            with pytest.raises(P.PrayError):
                ctx.render_pt_batch([sid], [cam], [(16, 16)], spp, depth, 3, [5], [bg], 1e-9)
    finally:
        ctx.free_scene(sid)
