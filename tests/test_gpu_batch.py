"""Batched multi-scene path tracing and the autotuner on the GPU: every scene
of a batch renders bit-identically to its single-scene render, and no tuning
changes a traced result."""
import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O
from conftest import first_diff, say

pytestmark = pytest.mark.gpu


def scene_set():
    """Random icosphere/box/terrain mixes of different sizes and cameras."""
    out = []
    for seed, tris, (w, h) in ((11, 3000, (64, 48)), (12, 20000, (96, 64)), (13, 1500, (32, 32)),
                               (14, 60000, (80, 80)), (15, 9000, (48, 96))):
        s = P.Scene.random(seed, tris)
        out.append((s, s.build(1), O.camera([0, 4, 9], [0, -0.45, -1], 60, w, h), (w, h), 1000 + seed))
    return out


def test_batch_equals_single_scene_renders(ctx):
    scenes = scene_set()
    sids = [ctx.upload(t) for _, t, _, _, _ in scenes]
    # one scene's tree built on the GPU (its own node layout) in the same batch
    abc, col, emit = scenes[1][0].triangles()
    ctx.free_scene(sids[1])
    sids[1], _ = ctx.build_scene(abc, col, emit, method=P.BUILD_SAH, leaf_size=8)
    try:
        single, single_rays = [], 0
        for k, (s, _, cam, (w, h), seed) in enumerate(scenes):
            say(f"single scene {k}")
            hdr, _ = ctx.render_pt(sids[k], cam, w, h, 4, 6, 3, seed, s.info()["background"], 1e-9, want_rgb=False)
            single_rays += ctx.last_render_rays()
            assert hdr.mean() > 0
            single.append(hdr)
        # The batch runs with the setting of scene 3 (the largest image of the
        # size bucket holding most pixels): the defaults, then 4-wave
        # workgroups with a 16-entry ring and two triangles per leaf step.
        for tun in (dict(), dict(wg_waves=4, stack_lds=16, leaf_ilp=2, fetch_min=24)):
            ctx.set_tuning(sids[3], **tun)
            say(f"batch of 5 scenes, 4 spp, scene 3 tuning {tun}")
            imgs, ms = ctx.render_pt_batch(sids, [c for _, _, c, _, _ in scenes],
                                           [wh for _, _, _, wh, _ in scenes], 4, 6, 3, [sd for *_, sd in scenes],
                                           [s.info()["background"] for s, *_ in scenes], 1e-9)
            assert ms > 0 and ctx.last_render_rays() == single_rays
            for k in range(len(scenes)):
                assert np.array_equal(imgs[k].view(np.uint32), single[k].view(np.uint32)), \
                    (k, tun, first_diff(imgs[k].ravel(), single[k].ravel()))
        # the passes enqueued kernel by kernel instead of replaying a graph
        ctx.use_graphs(False)
        try:
            say("batch of 5 scenes without graphs")
            imgs, ms = ctx.render_pt_batch(sids, [c for _, _, c, _, _ in scenes], [wh for _, _, _, wh, _ in scenes],
                                           4, 6, 3, [sd for *_, sd in scenes],
                                           [s.info()["background"] for s, *_ in scenes], 1e-9)
        finally:
            ctx.use_graphs(True)
        assert ms > 0 and ctx.last_render_rays() == single_rays
        for k in range(len(scenes)):
            assert np.array_equal(imgs[k].view(np.uint32), single[k].view(np.uint32)), k
    finally:
        for sid in sids:
            ctx.free_scene(sid)


@pytest.mark.parametrize("max_depth", [1, 0])
def test_batch_diffuse_estimator_equal_oracle(ctx, max_depth):
    """At a depth limit of at most 1 the batch takes the reference's diffuse
    estimator (pray.cpp:315-340) per scene, like the single-scene renderer:
    every scene's image equals the oracle's bit for bit."""
    scenes = scene_set()[:3]
    sids = [ctx.upload(t) for _, t, _, _, _ in scenes]
    try:
        say(f"batch of 3 scenes, diffuse estimator, max_depth {max_depth}")
        imgs, ms = ctx.render_pt_batch(sids, [c for _, _, c, _, _ in scenes], [wh for _, _, _, wh, _ in scenes],
                                       4, max_depth, 0, [sd for *_, sd in scenes],
                                       [s.info()["background"] for s, *_ in scenes], 1e-9)
    finally:
        for sid in sids:
            ctx.free_scene(sid)
    for k, (s, t, cam, (w, h), seed) in enumerate(scenes):
        oh, _ = O.render_pt(t.flat(), cam, w, h, 4, max_depth, 0, seed, s.info()["background"], 1e-9)
        assert np.array_equal(imgs[k].view(np.uint32), oh.view(np.uint32)), (k, first_diff(imgs[k].ravel(),
                                                                                           oh.ravel()))


def test_autotune_never_changes_hits(ctx):
    s = P.Scene.random(21, 40000)
    t = s.build(1)
    sid = ctx.upload(t)
    try:
        rays = O.gen_rays(5, P.scene_box_expanded(t.box()), 1 << 16)
        say("autotune: reference trace")
        ref_idx, ref_t = ctx.trace(sid, rays, P.TRACE_CLOSEST, 1e-9)
        say("autotune: tuning")
        tun = ctx.autotune(sid, n_rays=1 << 18, force=True)
        assert tun["stack_lds"] in (4, 6, 8, 16) and tun["leaf_ilp"] in (1, 2) and 1 <= tun["occupancy"] <= 100
        assert tun["fetch_min"] in (16, 24, 32, 48) and tun["wg_waves"] in (1, 2, 4)
        assert tun["leaf_batch"] in (0, 16, 24, 32)
        assert ctx.get_tuning(sid) == tun
        settings = [dict(), dict(stack_lds=4), dict(stack_lds=16, leaf_ilp=1), dict(fetch_min=16, occupancy=50),
                    dict(fetch_min=64, occupancy=75), dict(layout=P.LAYOUT_INDEXED, leaf_batch=0, node_levels=1),
                    dict(stack_lds=6),
                    dict(stack_lds=5, occupancy=10), dict(wg_waves=4), dict(wg_waves=2, stack_lds=16, leaf_ilp=2),
                    dict(layout=P.LAYOUT_PACKED, leaf_batch=16, stack_lds=8), dict(leaf_batch=48, occupancy=10),
                    dict(leaf_batch=1, wg_waves=4, fetch_min=1), dict(leaf_batch=0)]
        for st in settings:
            say(f"autotune: trace with {st}")
            ctx.set_tuning(sid, **st)
            idx, tt = ctx.trace(sid, rays, P.TRACE_CLOSEST, 1e-9)
            assert np.array_equal(idx, ref_idx) and np.array_equal(tt.view(np.uint32), ref_t.view(np.uint32)), st
        with pytest.raises(P.PrayError):
            ctx.set_tuning(sid, wg_waves=3)
        with pytest.raises(P.PrayError):
            ctx.set_tuning(sid, leaf_batch=65)
        with pytest.raises(P.PrayError):  # leaf batching traces the packed layout
            ctx.set_tuning(sid, layout=P.LAYOUT_INDEXED, leaf_batch=16)
        # a second scene of the same size bucket takes the bucket's setting
        s2 = P.Scene.random(22, 45000)
        sid2 = ctx.upload(s2.build(1))
        try:
            assert ctx.autotune(sid2) == tun
        finally:
            ctx.free_scene(sid2)
    finally:
        ctx.free_scene(sid)


def test_autotune_pt_keeps_images(ctx):
    """The path-tracing autotuner picks a valid setting for the scene alone
    and no image changes by a bit."""
    s = P.Scene.cornell(3, 20000)
    t = s.build(1)
    w, h, depth, rr = 64, 48, 6, 2
    cam = O.camera([0, 1, 3.8], [0, 0, -1], 40, w, h)
    bg = np.asarray(s.info()["background"], np.float32)
    sid = ctx.upload(t)
    sid2 = ctx.upload(t)
    try:
        ctx.set_tuning(sid, stack_lds=16, leaf_ilp=1, fetch_min=64, occupancy=50)
        say("autotune_pt: reference render")
        ref, _ = ctx.render_pt(sid, cam, w, h, 4, depth, rr, 5, bg, 1e-9, want_rgb=False)
        before2 = ctx.get_tuning(sid2)
        say("autotune_pt: tuning")
        tun = ctx.autotune_pt(sid, cam, w, h, depth, rr, 5, bg, 1e-9)
        assert tun["stack_lds"] in (4, 6, 8, 16) and tun["leaf_ilp"] in (1, 2)
        assert tun["fetch_min"] in (16, 24, 32, 48) and tun["occupancy"] in (50, 75, 100)
        assert tun["wg_waves"] in (1, 2, 4)
        assert ctx.get_tuning(sid) == tun
        assert ctx.get_tuning(sid2) == before2  # other scenes keep theirs
        hdr, _ = ctx.render_pt(sid, cam, w, h, 4, depth, rr, 5, bg, 1e-9, want_rgb=False)
        assert np.array_equal(hdr.view(np.uint32), ref.view(np.uint32))
        with pytest.raises(P.PrayError):
            ctx.autotune_pt(sid, cam, 0, h, depth, rr, 5, bg, 1e-9)
        with pytest.raises(P.PrayError):
            ctx.autotune_pt(12345, cam, w, h, depth, rr, 5, bg, 1e-9)
    finally:
        ctx.free_scene(sid)
        ctx.free_scene(sid2)
