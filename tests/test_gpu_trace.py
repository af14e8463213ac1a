"""GPU parity of the closest-hit / any-hit traversal (through the C ABI)
against the oracle: bit-exact triangle indices and distances."""
import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O
from conftest import first_diff, say

pytestmark = pytest.mark.gpu

SCENES = {}


def scene(name):
    """(scene, tree, flat) built once per session."""
    if name not in SCENES:
        if name == "spheres-median":
            s = P.Scene.spheres(1)
            t = s.build(0)
        elif name == "spheres-sah":
            s = P.Scene.spheres(1)
            t = s.build(1)
        elif name == "terrain256-sah":
            s = P.Scene.terrain(256, 2024)
            t = s.build(1)
        elif name == "terrain256-median":
            s = P.Scene.terrain(256, 2024)
            t = s.build(0)
        elif name == "cornell20k-sah":
            s = P.Scene.cornell(3, 20000)
            t = s.build(1)
        elif name == "random-median":
            s = P.Scene.random(5, 6000)
            t = s.build(0)
        else:
            raise KeyError(name)
        SCENES[name] = (s, t, t.flat())
    return SCENES[name]


# Every traversal setting gives the same hits:
# (layout, LDS ring entries, waves per workgroup, leaf ILP, refill threshold,
# leaf batch, tree levels per node step).  The first is the setting the
# autotuner picks for the benchmark's 2048x2048 terrain now (two levels per
# node step, ring 6, ILP 2, refill at 16 idle lanes); the next two are the
# ones it picked before two-level records (ring 8, ILP 1, refill at 24) and
# before that with leaf batching (ILP 2, refill at 16, leaf steps once 16
# lanes wait).  Two levels per step need a tree whose children pairs have
# equal clips (every SAH tree; the terrain's median tree too, whose children
# meet on grid lines); other trees refuse them.  Node records 3 (boxed
# children pairs) and 4 (their 4-wide form) run over the boxed layout the
# upload makes for every tree (the default for median trees whose children
# overlap; for SAH trees with the boxes cut to the KD cells).
SETTINGS = [(P.LAYOUT_PACKED, 6, 1, 2, 16, 0, 2),
            (P.LAYOUT_PACKED, 6, 1, 2, 16, 0, 3), (P.LAYOUT_PACKED, 8, 4, 1, 24, 16, 3),
            (P.LAYOUT_PACKED, 5, 1, 1, 1, 1, 3), (P.LAYOUT_PACKED, 16, 2, 2, 48, 0, 3),
            (P.LAYOUT_PACKED, 6, 1, 2, 16, 0, 4), (P.LAYOUT_PACKED, 4, 4, 1, 24, 16, 4),
            (P.LAYOUT_PACKED, 5, 1, 1, 1, 1, 4), (P.LAYOUT_PACKED, 16, 2, 2, 48, 0, 4),
            (P.LAYOUT_PACKED, 8, 1, 1, 24, 0, 1), (P.LAYOUT_PACKED, 8, 1, 2, 16, 16, 1),
            (P.LAYOUT_PACKED, 5, 2, 2, 16, 0, 1), (P.LAYOUT_INDEXED, 8, 4, 1, 32, 0, 1),
            (P.LAYOUT_INDEXED, 6, 1, 2, 16, 0, 1), (P.LAYOUT_PACKED, 6, 1, 2, 48, 24, 1),
            (P.LAYOUT_PACKED, 4, 2, 2, 16, 32, 1), (P.LAYOUT_PACKED, 16, 4, 1, 48, 64, 1),
            (P.LAYOUT_PACKED, 5, 1, 1, 1, 1, 1), (P.LAYOUT_PACKED, 8, 1, 2, 16, 16, 2),
            (P.LAYOUT_PACKED, 8, 1, 1, 24, 0, 2), (P.LAYOUT_PACKED, 4, 2, 2, 32, 32, 2),
            (P.LAYOUT_PACKED, 16, 4, 1, 1, 0, 2)]


def equal_clips(flat):
    """True when every children pair after pair 0 has two bitwise-equal
    clips, the condition for two-level node records."""
    nodes = flat["nodes"].reshape(-1, 2)
    if len(nodes) % 2:
        nodes = np.concatenate([nodes, np.zeros((1, 2), np.uint32)])
    pairs = nodes.reshape(-1, 2, 2)[1:]
    return bool(np.array_equal(pairs[:, 0, 0], pairs[:, 1, 0]))


def set_or_refuse(ctx, sid, name, layout, ring, wg, ilp, fm, lb, nl):
    """Applies a setting; False when the scene's tree has no two-level
    records (nl 2), after checking that the library refuses exactly those
    trees."""
    eq = equal_clips(scene(name)[2])
    if nl == 2 and not eq:
        with pytest.raises(P.PrayError):
            ctx.set_tuning(sid, layout=P.LAYOUT_PACKED, node_levels=nl)
        return False
    ctx.set_tuning(sid, layout=layout, stack_lds=ring, wg_waves=wg, leaf_ilp=ilp, fetch_min=fm, leaf_batch=lb,
                   node_levels=nl)
    return True


def check_closest(ctx, name, rays, eps, bf_rows=0):
    s, t, flat = scene(name)
    sid = ctx.upload(t)
    try:
        for layout, ring, wg, ilp, fm, lb, nl in SETTINGS:
            if not set_or_refuse(ctx, sid, name, layout, ring, wg, ilp, fm, lb, nl):
                continue
            say(f"closest {name} layout={layout} stack_lds={ring} wg={wg} ilp={ilp} fetch_min={fm} leaf_batch={lb} "
                f"node_levels={nl} n={len(rays)} eps={eps}")
            gi, gt = ctx.trace(sid, rays, P.TRACE_CLOSEST, eps)
            oi, ot = O.flat_trace(flat, rays, 0, eps)
            assert first_diff(gi, oi) is None, first_diff(gi, oi)
            assert np.array_equal(gt.view(np.uint32), ot.view(np.uint32))
        if bf_rows:
            sub = rays[:bf_rows]
            bi, bt = O.bf_closest(flat["a"], flat["ab"], flat["ac"], sub, eps)
            assert first_diff(gi[:bf_rows], bi) is None, first_diff(gi[:bf_rows], bi)
            h = bi >= 0
            assert np.array_equal(gt[:bf_rows][h], bt[h])
        return gi, gt
    finally:
        ctx.free_scene(sid)


@pytest.mark.parametrize("name", ["spheres-median", "spheres-sah", "cornell20k-sah"])
def test_closest_camera_rays(ctx, name):
    s, _, _ = scene(name)
    rays = s.prim_rays()
    gi, _ = check_closest(ctx, name, rays, 1e-4, bf_rows=4096)
    assert (gi >= 0).mean() > 0.2


@pytest.mark.parametrize("name,eps", [("spheres-median", 1e-4), ("spheres-sah", 1e-4), ("terrain256-sah", 1e-12),
                                      ("terrain256-sah", 1e-4), ("terrain256-median", 1e-12),
                                      ("cornell20k-sah", 1e-9)])
def test_closest_incoherent(ctx, name, eps):
    _, _, flat = scene(name)
    rays = O.gen_rays(77, P.scene_box_expanded(flat["box"]), 1 << 16)
    gi, gt = check_closest(ctx, name, rays, eps, bf_rows=2048)
    if name.startswith("terrain") and eps == 1e-4:
        # every terrain triangle is culled by the reference's absolute epsilon
        assert (gi < 0).all()


@pytest.mark.parametrize("mode", [P.TRACE_ANY, P.TRACE_ANY_INCLUSIVE])
@pytest.mark.parametrize("name,eps", [("spheres-median", 1e-4), ("terrain256-sah", 1e-12), ("cornell20k-sah", 1e-9)])
def test_any_hit(ctx, name, eps, mode):
    s, t, flat = scene(name)
    rays = O.gen_rays(91, P.scene_box_expanded(flat["box"]), 1 << 15)
    ci, ct = O.flat_trace(flat, rays, 0, eps)
    rng = np.random.default_rng(5)
    np.seterr(over="ignore")  # t*t of misses (FLT_MAX) overflows; those lanes take the other branch
    # limits around the closest hit distance, some exactly equal to t*t
    lim = np.where(ci >= 0, (ct * ct) * rng.choice([0.5, 1.0, 2.0], size=len(ct)).astype(np.float32),
                   rng.uniform(0, 4, len(ct)).astype(np.float32)).astype(np.float32)
    sid = ctx.upload(t)
    try:
        res = []
        for layout, ring, wg, ilp, fm, lb, nl in SETTINGS:
            if not set_or_refuse(ctx, sid, name, layout, ring, wg, ilp, fm, lb, nl):
                continue
            say(f"any {name} mode={mode} layout={layout} stack_lds={ring} wg={wg} ilp={ilp} fetch_min={fm} "
                f"leaf_batch={lb} node_levels={nl}")
            res.append(ctx.trace(sid, rays, mode, eps, limit2=lim))
    finally:
        ctx.free_scene(sid)
    sq = ct * ct
    want = (ci >= 0) & ((sq < lim) if mode == P.TRACE_ANY else (sq <= lim))
    for gi, _ in res:
        assert np.array_equal(gi >= 0, want)
    gi, gt = res[0]
    occ = np.flatnonzero(gi >= 0)
    # every reported occluder is a genuine hit under the same predicate
    for k in occ[:512]:
        h, tt = O.mt(rays[k], flat["a"][gi[k]], flat["ab"][gi[k]], flat["ac"][gi[k]], eps)
        assert h and tt > 0 and tt == gt[k]
        assert (tt * np.float32(tt) < lim[k]) if mode == P.TRACE_ANY else (np.float32(tt) * np.float32(tt) <= lim[k])


@pytest.mark.parametrize("lb,nl", [(0, 1), (16, 1), (40, 1), (0, 2), (16, 2)])
def test_work_counts_match_cpu_twin(ctx, lb, nl):
    """Per-ray node and triangle counts equal the CPU twin's, with and
    without leaf batching, one or two tree levels per node step."""
    _, t, flat = scene("terrain256-sah")
    rays = O.gen_rays(3, P.scene_box_expanded(flat["box"]), 1 << 14)
    sid = ctx.upload(t)
    try:
        ctx.set_tuning(sid, layout=P.LAYOUT_PACKED, leaf_batch=lb, node_levels=nl)
        say(f"counts terrain256-sah leaf_batch={lb} node_levels={nl}")
        n = rays.shape[0]
        d_rays = ctx.to_device(rays)
        d_idx = ctx.alloc(n, np.int32)
        d_t = ctx.alloc(n, np.float32)
        d_c = ctx.alloc((n, 2), np.uint32)
        ctx.trace_counts_device(sid, d_rays.ptr, n, d_idx.ptr, d_t.ptr, d_c.ptr, P.TRACE_CLOSEST, 1e-12)
        ctx.sync()
        gc = d_c.numpy()
        gi = d_idx.numpy()
    finally:
        ctx.free_scene(sid)
    oi, ot, oc = O.flat_trace(flat, rays, 0, 1e-12, counts=True)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gc, oc)


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("lb", [0, 16])
@pytest.mark.parametrize("name,rays_of", [("spheres-median", "camera"), ("spheres-median", "incoherent"),
                                          ("random-median", "incoherent"), ("spheres-sah", "camera"),
                                          ("terrain256-sah", "incoherent"), ("cornell20k-sah", "incoherent")])
def test_work_counts_boxed_match_cpu_twin(ctx, name, rays_of, lb, width):
    """Boxed node steps (node records 3, and 4 over the 4-wide form, the
    default for median trees whose children overlap; over SAH trees with the
    boxes cut to the KD cells): per-ray node and triangle counts and hits
    equal the CPU twin's over the same boxed layout, with and without leaf
    batching; the twin's hits equal the plane-form traversal's and brute
    force."""
    s, t, flat = scene(name)
    rays = s.prim_rays() if rays_of == "camera" else O.gen_rays(4, P.scene_box_expanded(flat["box"]), 1 << 15)
    sid = ctx.upload(t)
    try:
        assert ctx.get_tuning(sid)["node_levels"] == 4
        ctx.set_tuning(sid, leaf_batch=lb, node_levels=3 if width == 2 else 4)
        say(f"counts boxed {name} {rays_of} leaf_batch={lb} width={width}")
        n = rays.shape[0]
        d_rays = ctx.to_device(rays)
        d_idx = ctx.alloc(n, np.int32)
        d_t = ctx.alloc(n, np.float32)
        d_c = ctx.alloc((n, 2), np.uint32)
        ctx.trace_counts_device(sid, d_rays.ptr, n, d_idx.ptr, d_t.ptr, d_c.ptr, P.TRACE_CLOSEST, 1e-4)
        ctx.sync()
        gc, gi, gt = d_c.numpy(), d_idx.numpy(), d_t.numpy()
    finally:
        ctx.free_scene(sid)
    oi, ot, oc = O.flat_trace(flat, rays, 0, 1e-4, counts=True, boxed=P.boxed_layout(flat, width))
    assert np.array_equal(gi, oi)
    assert np.array_equal(gt.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(gc, oc)
    pi, pt, pc = O.flat_trace(flat, rays, 0, 1e-4, counts=True)
    assert np.array_equal(oi, pi) and np.array_equal(ot.view(np.uint32), pt.view(np.uint32))
    # the boxes prune: fewer triangle tests than the clip planes
    assert oc[:, 1].sum() < pc[:, 1].sum()
    bi, bt = O.bf_closest(flat["a"], flat["ab"], flat["ac"], rays[:2048], 1e-4)
    assert first_diff(oi[:2048], bi) is None


def test_node_records_by_tree(ctx):
    """The upload's default node records: the 4-wide boxed layout for every
    tree (the median tree whose children overlap, SAH trees, the terrain's
    median tree with equal clips); every uploaded tree takes the other
    forms when asked, two-level records only where clips are equal."""
    for name in ("spheres-median", "spheres-sah", "terrain256-median"):
        _, t, flat = scene(name)
        sid = ctx.upload(t)
        try:
            assert ctx.get_tuning(sid)["node_levels"] == 4, name
            for nl in (3, 4, 1) + ((2,) if equal_clips(flat) else ()):
                ctx.set_tuning(sid, node_levels=nl)
                assert ctx.get_tuning(sid)["node_levels"] == nl
            if not equal_clips(flat):
                with pytest.raises(P.PrayError):
                    ctx.set_tuning(sid, node_levels=2)
        finally:
            ctx.free_scene(sid)


@pytest.mark.parametrize("name,ring,lb,nl", [("spheres-sah", 5, 0, 1), ("spheres-sah", 16, 0, 1),
                                             ("spheres-sah", 5, 32, 1), ("spheres-sah", 16, 8, 1),
                                             ("spheres-sah", 5, 0, 2), ("spheres-sah", 16, 24, 2),
                                             ("spheres-median", 5, 0, 3), ("spheres-median", 16, 24, 3),
                                             ("spheres-median", 4, 0, 4), ("spheres-median", 16, 24, 4),
                                             ("spheres-sah", 8, 0, 4), ("spheres-sah", 5, 16, 3)])
def test_launch_sizes(ctx, name, ring, lb, nl):
    """Traces over ray counts around the wave, the 64-ray chunks and the 64
    chunk counters (partial last chunks, fewer waves than counters), one after
    another through the same counters and spill area, including a ray without
    direction (its hit is not compared) and one far outside the scene box.
    The last launch has a grid of a few hundred waves for 3125 chunks, so the
    waves walk the counters as they run dry."""
    _, t, flat = scene(name)
    sid = ctx.upload(t)
    try:
        ctx.set_tuning(sid, layout=P.LAYOUT_PACKED, stack_lds=ring, fetch_min=1 if ring == 5 else 64,
                       wg_waves=1 if ring == 5 else 4, leaf_batch=lb, node_levels=nl)
        base = O.gen_rays(12, P.scene_box_expanded(flat["box"]), 200_000)
        base[5, 3:] = 0.0
        base[7, :3] = [1e30, -1e30, 0.0]
        keep = np.arange(len(base)) != 5
        for n in (1, 63, 64, 65, 4095, 4096, 4097, 9000, 100, 12293, 200_000):
            if n == 200_000:
                ctx.set_tuning(sid, occupancy=1)
            rays = np.ascontiguousarray(base[:n])
            say(f"launch ring={ring} leaf_batch={lb} node_levels={nl} n={n}")
            gi, gt = ctx.trace(sid, rays, P.TRACE_CLOSEST, 1e-4)
            oi, ot = O.flat_trace(flat, rays, 0, 1e-4)
            k = keep[:n]
            assert first_diff(gi[k], oi[k]) is None, (n, first_diff(gi[k], oi[k]))
            assert np.array_equal(gt[k].view(np.uint32), ot[k].view(np.uint32))
    finally:
        ctx.free_scene(sid)


def test_gen_rays_bitwise(ctx):
    box = np.array([-0.05, -0.01, -0.05, 1.1, 0.12, 1.1], np.float32)
    say("gen rays")
    g = ctx.gen_rays(123456789, box, 100_000)
    o = O.gen_rays(123456789, box, 100_000)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    d = g[:, 3:].astype(np.float64)
    assert np.allclose((d * d).sum(1), 1.0, atol=1e-6)


def test_camera_rays_bitwise(ctx):
    s, _, _ = scene("spheres-median")
    cam = s.camera()
    say("camera rays")
    g = ctx.camera_rays(cam, 512, 512)
    assert np.array_equal(g.view(np.uint32), s.prim_rays().view(np.uint32))
    if O.have_ref():
        r = O.ref_prim_rays([0, 3, 8], [0, -0.35, -1], 60, 512, 512)
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


def test_empty_scene(ctx):
    s = P.Scene.from_triangles(np.zeros((0, 9), np.float32))
    t = s.build(0)
    sid = ctx.upload(t)
    try:
        rays = O.gen_rays(1, np.array([0, 0, 0, 1, 1, 1], np.float32), 1000)
        gi, gt = ctx.trace(sid, rays, P.TRACE_CLOSEST, 1e-4)
        assert (gi == -1).all() and (gt == np.finfo(np.float32).max).all()
    finally:
        ctx.free_scene(sid)


def test_axis_parallel_and_degenerate_rays(ctx):
    """Rays with zero direction components and origins on split planes and
    on the scene box faces (the slab/NaN cases of the interval arithmetic)."""
    _, t, flat = scene("terrain256-sah")
    rng = np.random.default_rng(8)
    n = 20000
    o = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    o[:, 1] = rng.choice([0.2, 0.05, 0.0, 0.1], n).astype(np.float32)
    grid = (rng.integers(0, 257, n) / 256).astype(np.float32)
    o[: n // 2, 0] = grid[: n // 2]  # exactly on cell boundaries (split candidates)
    d = np.zeros((n, 3), np.float32)
    axis = rng.integers(0, 3, n)
    d[np.arange(n), axis] = rng.choice([-1.0, 1.0], n)
    d[n // 3:, 1] = -1.0  # straight down
    d[n // 4: n // 3] = rng.normal(size=(n // 3 - n // 4, 3)).astype(np.float32)
    d[n // 4: n // 3, 0] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    d = d.astype(np.float32)
    d[d == 0] = rng.choice([0.0, -0.0], int((d == 0).sum())).astype(np.float32)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    check_closest(ctx, "terrain256-sah", rays, 1e-12, bf_rows=3000)


def test_kernel_timing_ring(ctx):
    """pray_timed_kernels_ms averages the traversal kernels enqueued back to
    back since timing was switched on (the last 64 of them); the per-kernel
    mean fits inside the events around the whole region, and timing does not
    change the hits."""
    _, t, flat = scene("terrain256-sah")
    sid = ctx.upload(t, shading=False)
    try:
        rays = O.gen_rays(3, P.scene_box_expanded(t.box()), 1 << 16)
        d_rays = ctx.to_device(rays)
        d_idx, d_t = ctx.alloc(len(rays), np.int32), ctx.alloc(len(rays), np.float32)
        ctx.trace_device(sid, d_rays.ptr, len(rays), d_idx.ptr, d_t.ptr, P.TRACE_CLOSEST, 1e-12)
        want = d_idx.numpy()
        with pytest.raises(P.PrayError):
            ctx.kernel_timing(True)
            ctx.timed_kernels_ms()
        for launches, expect in ((3, 3), (70, 64)):
            ctx.kernel_timing(False)
            ctx.kernel_timing(True)
            ctx.record(0)
            for _ in range(launches):
                ctx.trace_device(sid, d_rays.ptr, len(rays), d_idx.ptr, d_t.ptr, P.TRACE_CLOSEST, 1e-12)
            ctx.record(1)
            mean, cnt = ctx.timed_kernels_ms()
            region = ctx.elapsed_ms(0, 1)
            assert cnt == expect and 0 < mean and mean * cnt <= region, (launches, mean, cnt, region)
            assert 0 < ctx.last_kernel_ms() <= region
            assert np.array_equal(d_idx.numpy(), want)
        ctx.kernel_timing(False)
        for b in (d_rays, d_idx, d_t):
            b.free()
    finally:
        ctx.free_scene(sid)
