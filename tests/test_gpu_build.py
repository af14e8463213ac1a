"""The GPU tree build (pray_build_scene) against the oracle's twin: the
device layout word for word, closest hits equal to brute force, and path
tracing identical to the CPU-built tree's (closest hits do not depend on the
tree)."""
import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O
from conftest import first_diff, say
from test_oracle_build import SCENES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("leaf_size", [1, 8])
@pytest.mark.parametrize("name", sorted(SCENES))
@pytest.mark.parametrize("method", [P.BUILD_MORTON, P.BUILD_SAH])
def test_gpu_build_equals_oracle(ctx, method, name, leaf_size):
    abc = np.ascontiguousarray(SCENES[name](), np.float32).reshape(-1, 9)
    n = abc.shape[0]
    say(f"gpu build method={method} {name} n={n} leaf={leaf_size}")
    sid, ms = ctx.build_scene(abc, method=method, leaf_size=leaf_size, max_depth=24 if leaf_size == 1 else 40)
    try:
        g = ctx.scene_layout(sid)
        if method == P.BUILD_MORTON:
            o = O.lbvh_build(abc, leaf_size)
        else:
            o = O.sah_build(abc, leaf_size, 24 if leaf_size == 1 else 40, 2.0, 1.0)
        assert g["n_tris"] == n and g["max_depth"] == o["max_depth"]
        assert np.array_equal(g["box"].view(np.uint32), o["box"].view(np.uint32))
        assert np.array_equal(g["refs"], o["refs"])
        assert first_diff(g["nodes"], o["nodes"]) is None, first_diff(g["nodes"], o["nodes"])
        # the boxed layout is made on the host at upload, so a tree built on
        # the device has none
        for nl in (3, 4):
            with pytest.raises(P.PrayError):
                ctx.set_tuning(sid, node_levels=nl)
        rays = O.gen_rays(23, P.scene_box_expanded(o["box"]), 5000)
        say(f"gpu build trace {name}")
        gi, gt = ctx.trace(sid, rays, P.TRACE_CLOSEST, 1e-9)
        bi, bt = O.bf_closest(o["a"], o["ab"], o["ac"], rays, 1e-9)
        assert np.array_equal(gi, bi)
        h = bi >= 0
        assert np.array_equal(gt[h].view(np.uint32), bt[h].view(np.uint32))
    finally:
        ctx.free_scene(sid)


@pytest.mark.parametrize("method", [P.BUILD_MORTON, P.BUILD_SAH])
def test_gpu_build_empty_scene(ctx, method):
    sid, _ = ctx.build_scene(np.zeros((0, 9), np.float32), method=method)
    try:
        rays = O.gen_rays(1, np.array([0, 0, 0, 1, 1, 1], np.float32), 100)
        gi, gt = ctx.trace(sid, rays, P.TRACE_CLOSEST, 1e-4)
        assert (gi == -1).all()
    finally:
        ctx.free_scene(sid)


def test_gpu_build_rejects_bad_leaf_size(ctx):
    with pytest.raises(P.PrayError):
        ctx.build_scene(np.zeros((4, 9), np.float32), leaf_size=0)
    with pytest.raises(P.PrayError):
        ctx.build_scene(np.zeros((4, 9), np.float32), leaf_size=65)
    with pytest.raises(P.PrayError):
        ctx.build_scene(np.zeros((4, 9), np.float32), max_depth=65)
    with pytest.raises(P.PrayError):
        ctx.build_scene(np.zeros((4, 9), np.float32), method=7)


def test_gpu_built_scene_path_traces_like_cpu_tree(ctx):
    """Shading data computed on the device equals the host precompute, and
    the image equals the CPU SAH tree's and the oracle's bit for bit."""
    s = P.Scene.cornell(3, 20000)
    abc, col, emit = s.triangles()
    w, h, spp, depth, rr = 64, 48, 4, 6, 2
    cam = O.camera([0, 1, 3.8], [0, 0, -1], 40, w, h)
    bg = np.asarray(s.info()["background"], np.float32)
    t = s.build(1)
    say("pt on the cpu-built tree")
    sid = ctx.upload(t)
    try:
        ref_hdr, ref_rgb = ctx.render_pt(sid, cam, w, h, spp, depth, rr, 9, bg, 1e-9)
    finally:
        ctx.free_scene(sid)
    for method in (P.BUILD_MORTON, P.BUILD_SAH):
        say(f"pt on the gpu-built tree, method {method}")
        gid, _ = ctx.build_scene(abc, col, emit, method=method, leaf_size=4)
        try:
            hdr, rgb = ctx.render_pt(gid, cam, w, h, spp, depth, rr, 9, bg, 1e-9)
        finally:
            ctx.free_scene(gid)
        assert np.array_equal(hdr.view(np.uint32), ref_hdr.view(np.uint32)) and np.array_equal(rgb, ref_rgb)
    oh, _ = O.render_pt(t.flat(), cam, w, h, spp, depth, rr, 9, bg, 1e-9)
    assert np.array_equal(hdr.view(np.uint32), oh.view(np.uint32))
    assert hdr.mean() > 0.01


def test_gpu_build_terrain_full_size(ctx):
    """The 2048 x 2048-cell terrain (8,388,608 triangles), both methods:
    every triangle in a leaf, and 2^20 incoherent rays hit exactly as over
    the CPU SAH tree."""
    s = P.Scene.terrain(2048, 2024)
    abc, _, _ = s.triangles()
    box = None
    res = {}
    for method in (P.BUILD_MORTON, P.BUILD_SAH):
        say(f"gpu build terrain2048 method {method}")
        gid, ms = ctx.build_scene(abc, method=method, leaf_size=8)
        print(f"GPU build (method {method}) of {abc.shape[0]} triangles: {ms:.2f} ms", flush=True)
        try:
            g = ctx.scene_layout(gid)
            assert g["max_depth"] <= 64
            tri = np.unique(g["refs"] & 0x7FFFFFFF)
            assert np.array_equal(tri, np.arange(abc.shape[0], dtype=np.uint32))
            if method == P.BUILD_MORTON:
                # word for word: some levels hold more inner nodes (at least
                # 398,870 at depth 24) than one pass of the level-order
                # kernel's 1024 x 256 threads
                say("oracle twin of the terrain2048 Morton build")
                o = O.lbvh_build(np.ascontiguousarray(abc, np.float32).reshape(-1, 9), 8)
                assert g["max_depth"] == o["max_depth"]
                assert np.array_equal(g["box"].view(np.uint32), o["box"].view(np.uint32))
                assert np.array_equal(g["refs"], o["refs"])
                assert first_diff(g["nodes"], o["nodes"]) is None, first_diff(g["nodes"], o["nodes"])
                del o
            if box is None:
                box = g["box"]
                rays = O.gen_rays(2024, P.scene_box_expanded(box), 1 << 20)
            say(f"trace terrain2048 gpu-built method {method}")
            res[method] = ctx.trace(gid, rays, P.TRACE_CLOSEST, 1e-12)
        finally:
            ctx.free_scene(gid)
    t = s.build(1, max_depth=40, leaf_size=8, traversal_cost=2.0)
    sid = ctx.upload(t, shading=False)
    try:
        say("trace terrain2048 cpu-built")
        ci, ct = ctx.trace(sid, rays, P.TRACE_CLOSEST, 1e-12)
    finally:
        ctx.free_scene(sid)
    for gi, gt in res.values():
        assert np.array_equal(gi, ci) and np.array_equal(gt.view(np.uint32), ct.view(np.uint32))
    assert (ci >= 0).mean() > 0.1


@pytest.mark.parametrize("name", ["spheres", "terrain96", "cornell", "random"])
def test_gpu_sah_build_equals_cpu_binned_tree(ctx, name):
    """The GPU SAH build and the host's binned-SAH builder make the same tree
    (same splits and leaf lists, different node order)."""
    from test_oracle_build import canonical
    abc = np.ascontiguousarray(SCENES[name](), np.float32).reshape(-1, 9)
    say(f"gpu sah vs cpu binned {name}")
    sid, _ = ctx.build_scene(abc, method=P.BUILD_SAH, leaf_size=8, max_depth=40)
    try:
        g = ctx.scene_layout(sid)
    finally:
        ctx.free_scene(sid)
    f = P.Scene.from_triangles(abc).build(2, max_depth=40, leaf_size=8, traversal_cost=2.0).flat()
    assert canonical(g) == canonical(f)


def test_gpu_build_reuses_scratch(ctx):
    """Builds on one context keep their scratch buffers: a small scene built
    before and after a larger one (whose data stays in those buffers) gives
    the same layout word for word, for both methods."""
    small = np.ascontiguousarray(P.Scene.random(4, 6000).triangles()[0], np.float32).reshape(-1, 9)
    large = np.ascontiguousarray(P.Scene.terrain(256, 5).triangles()[0], np.float32).reshape(-1, 9)
    for method in (P.BUILD_MORTON, P.BUILD_SAH):
        layouts = []
        for abc in (small, large, small):
            say(f"gpu build reuse method={method} n={abc.shape[0]}")
            sid, _ = ctx.build_scene(abc, method=method, leaf_size=8)
            try:
                layouts.append(ctx.scene_layout(sid))
            finally:
                ctx.free_scene(sid)
        a, b = layouts[0], layouts[2]
        assert a["max_depth"] == b["max_depth"] and a["n_tris"] == small.shape[0]
        assert np.array_equal(a["box"].view(np.uint32), b["box"].view(np.uint32))
        assert np.array_equal(a["refs"], b["refs"])
        assert first_diff(a["nodes"], b["nodes"]) is None, first_diff(a["nodes"], b["nodes"])
