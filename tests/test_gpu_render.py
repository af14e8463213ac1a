"""GPU renderers against the oracle and against images rendered by the
reference compiled from source (tests/golden/ref_images.npz)."""
import os
import subprocess

import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O
from conftest import first_diff, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
RT_SCENES = ["cube_200x150.json", "cube_203x101.json", "head_fov30.json", "head_fov30_odd.json", "plant_256.json"]


def load(name):
    s = P.Scene.load_json(os.path.join(GOLDEN, "scenes", name))
    t = s.build(-1)
    return s, t, t.flat()


@pytest.mark.parametrize("name", RT_SCENES)
def test_rt_example_scenes_equal_reference_images(ctx, name):
    s, t, f = load(name)
    inf = s.info()
    sid = ctx.upload(t)
    try:
        say(f"rt {name}")
        img, hits = ctx.render_rt(sid, s.camera(), inf["width"], inf["height"], s.lights(), inf["background"],
                                  P.REFERENCE_DET_EPS, want_hits=True)
    finally:
        ctx.free_scene(sid)
    ref = np.load(os.path.join(GOLDEN, "ref_images.npz"))[f"{name}:exact"]
    assert img.shape == ref.shape
    assert np.array_equal(img, ref), f"{np.count_nonzero(np.any(img != ref, axis=2))} pixels differ"
    # the FMA build of the reference differs from its own exact build on a
    # few shadow-boundary pixels; report-only
    fma = np.load(os.path.join(GOLDEN, "ref_images.npz"))[f"{name}:fma"]
    print("pixels differing from the reference FMA build:", int(np.count_nonzero(np.any(img != fma, axis=2))))
    # primary hits: equal to the reference KD tree except at equal-t ties
    g = np.load(os.path.join(GOLDEN, "ref_hits.npz"))
    ri, rt = g[f"{name}:idx"], g[f"{name}:t"]
    diff = np.flatnonzero(hits != ri)
    assert diff.size <= max(2, hits.size // 1000)
    rays = s.prim_rays()
    for k in diff:  # a tie: both triangles are hit at the same t
        h1, t1 = O.mt(rays[k], f["a"][hits[k]], f["ab"][hits[k]], f["ac"][hits[k]], 1e-4)
        assert h1 and np.float32(t1) == rt[k] and hits[k] < ri[k]


@pytest.mark.parametrize("seed", [1, 2])
def test_rt_spheres_equal_oracle(ctx, seed):
    s = P.Scene.spheres(seed)
    t = s.build(0)
    f = t.flat()
    inf = s.info()
    sid = ctx.upload(t)
    try:
        say(f"rt spheres {seed}")
        img = ctx.render_rt(sid, s.camera(), inf["width"], inf["height"], s.lights(), inf["background"],
                            P.REFERENCE_DET_EPS)
        # a frame loop's reused buffers (page-locked): filled over whatever
        # they held
        out = (ctx.host_array((inf["height"], inf["width"], 3), np.uint8),
               ctx.host_array(inf["width"] * inf["height"], np.int32))
        out[0][:] = 7
        out[1][:] = 5
        for _ in range(2):
            img2, hits2 = ctx.render_rt(sid, s.camera(), inf["width"], inf["height"], s.lights(), inf["background"],
                                        P.REFERENCE_DET_EPS, want_hits=True, out=out)
            assert img2 is out[0] and hits2 is out[1]
            assert np.array_equal(img2, img)
        with pytest.raises(ValueError):
            ctx.render_rt(sid, s.camera(), inf["width"], inf["height"], s.lights(), inf["background"],
                          P.REFERENCE_DET_EPS, want_hits=True, out=(out[0][:-1], out[1]))
    finally:
        ctx.free_scene(sid)
    ref = O.render_rt(f, s.camera(), inf["width"], inf["height"], s.lights(), inf["background"], 1e-4)
    assert np.array_equal(img, ref), f"{np.count_nonzero(np.any(img != ref, axis=2))} pixels differ"
    assert img.std() > 10  # not a blank image
    rays = s.prim_rays()
    bi, _ = O.bf_closest(f["a"], f["ab"], f["ac"], rays[::97], 1e-4)
    assert np.array_equal(hits2[::97], bi)


def test_rt_multi_light_negative_lambert_quirk(ctx):
    """Two lights on opposite sides: back-facing terms enter packets where a
    neighbour faces the light (pray.cpp:129); checked against the oracle."""
    s = P.Scene.spheres(4)
    t = s.build(0)
    f = t.flat()
    inf = s.info()
    lights = np.array([[4, 8, 6, 60, 60, 60], [-6, 2, -3, 40, 10, 10], [0, 0.5, 4, 5, 30, 5]], np.float32)
    sid = ctx.upload(t)
    try:
        # widths below 4 are undefined in the reference (pray.cpp:388 loops
        # to width - 4 in uint64_t); the smallest packet image is 4 x 2.
        # Frames of shrinking, then growing sizes reuse the context's
        # workspace buffers.
        for w, h in [(512, 512), (101, 77), (6, 3), (4, 2), (203, 101)]:
            say(f"rt multi-light {w}x{h}")
            cam = O.camera([0, 3, 8], [0, -0.35, -1], 60, w, h)
            img = ctx.render_rt(sid, cam, w, h, lights, inf["background"], 1e-4)
            ref = O.render_rt(f, cam, w, h, lights, inf["background"], 1e-4)
            assert np.array_equal(img, ref), (w, h)
    finally:
        ctx.free_scene(sid)


# (hipGraph replays, path budget): all 8 samples in one pass (the default
# budget), passes of 3, 3 and 2 samples, and one sample per pass; with and
# without graphs.
# The bounces take the upload's default 4-wide boxed node steps, over the
# room's SAH tree and (the last two) its median tree; the SAH cases without
# graphs take the KD tree's plane-form node steps instead.
@pytest.mark.parametrize("use_graph,budget,kind", [(1, None, 1), (0, None, 1), (1, 3 * 64 * 48, 1),
                                                   (0, 3 * 64 * 48, 1), (1, 0, 1), (0, 0, 1), (1, None, 0),
                                                   (0, 3 * 64 * 48, 0)])
def test_pt_cornell_equal_oracle(ctx, use_graph, budget, kind):
    s = P.Scene.cornell(3, 20000)
    t = s.build(kind)
    f = t.flat()
    inf = s.info()
    w, h, spp, depth, rr = 64, 48, 8, 6, 2
    cam = O.camera([0, 1, 3.8], [0, 0, -1], 40, w, h)
    sid = ctx.upload(t)
    assert ctx.get_tuning(sid)["node_levels"] == 4
    if kind == 1 and not use_graph:
        ctx.set_tuning(sid, node_levels=1)
    if budget is not None:
        ctx.set_path_budget(budget)
    try:
        say(f"pt cornell graph={use_graph} budget={budget} tree={kind}")
        hdr = np.empty((h, w, 3), np.float32)
        rgb = np.empty((h, w, 3), np.uint8)
        lib = P.libraries()[0]
        lib.pray_render_pt_ex.restype = P.C.c_int
        lib.pray_render_pt_ex.argtypes = [P._vp, P.C.c_int, P._vp, P._u32, P._u32, P._u32, P._u32, P._u32, P._u64,
                                          P._vp, P._f32, P._vp, P._vp, P.C.c_int]
        bg = np.asarray(inf["background"], np.float32)
        ctx._check(lib.pray_render_pt_ex(ctx.handle, sid, P._p(cam), w, h, spp, depth, rr, 77, P._p(bg), 1e-9,
                                         P._p(hdr), P._p(rgb), use_graph))
    finally:
        ctx.set_path_budget(1 << 24)
        ctx.free_scene(sid)
    queues = ctx.last_render_queues()
    oh, orgb, oq = O.render_pt(f, cam, w, h, spp, depth, rr, 77, bg, 1e-9, queues=True)
    assert np.array_equal(hdr.view(np.uint32), oh.view(np.uint32))
    assert np.array_equal(rgb, orgb)
    assert hdr.mean() > 0.01
    # the wavefront queues hold exactly the paths the per-pixel restatement
    # traces at each bounce (all samples), and they shrink
    assert np.array_equal(queues, oq), (queues, oq)
    assert queues[0] == w * h * spp and (np.diff(queues.astype(np.int64)) <= 0).all() and queues[-1] > 0


# The reference's diffuse estimator at a depth limit of at most 1
# (pray.cpp:315-340): graphs, passes of 3 + 3 + 2 samples, both trees, and a
# progressive split into 3 + 5 samples.
@pytest.mark.parametrize("max_depth,use_graph,budget,kind", [(1, 1, None, 1), (1, 0, 3 * 64 * 48, 1), (1, 1, None, 0),
                                                             (0, 1, None, 1), (0, 0, 3 * 64 * 48, 0)])
def test_pt_diffuse_estimator_equal_oracle(ctx, max_depth, use_graph, budget, kind):
    s = P.Scene.cornell(3, 20000)
    t = s.build(kind)
    f = t.flat()
    w, h, spp = 64, 48, 8
    cam = O.camera([0, 1, 3.8], [0, 0, -1], 40, w, h)
    bg = np.asarray(s.info()["background"], np.float32)
    sid = ctx.upload(t)
    if budget is not None:
        ctx.set_path_budget(budget)
    try:
        say(f"pt diffuse max_depth={max_depth} graph={use_graph} budget={budget} tree={kind}")
        hdr, rgb = ctx.render_pt(sid, cam, w, h, spp, max_depth, 0, 77, bg, 1e-9, use_graph=bool(use_graph))
        queues = ctx.last_render_queues()
        acc, first = None, 0
        for n in (3, 5):
            ph, prgb, acc = ctx.render_pt_progressive(sid, cam, w, h, first, n, max_depth, 0, 77, bg, 1e-9,
                                                      accum=acc)
            first += n
    finally:
        ctx.set_path_budget(1 << 24)
        ctx.free_scene(sid)
    oh, orgb, oq = O.render_pt(f, cam, w, h, spp, max_depth, 0, 77, bg, 1e-9, queues=True)
    assert np.array_equal(hdr.view(np.uint32), oh.view(np.uint32)), first_diff(hdr.ravel(), oh.ravel())
    assert np.array_equal(rgb, orgb)
    assert np.array_equal(queues, oq), (queues, oq)
    assert np.array_equal(ph.view(np.uint32), hdr.view(np.uint32)) and np.array_equal(prgb, rgb)
    # the pixels the diffuse form fixes: background on a miss and emission
    # on an emissive hit, exactly, not averaged over the samples
    hits, _ = O.flat_trace(f, O.camera_rays(cam, w, h), 0, 1e-9)
    miss = hits < 0
    em = (hits >= 0) & (f["emit"][np.maximum(hits, 0)] != 0)
    img = hdr.reshape(-1, 3)
    assert np.array_equal(img[miss], np.broadcast_to(bg, (int(miss.sum()), 3)))
    assert em.any() and np.array_equal(img[em], f["color"].reshape(-1, 3)[hits[em]])
    if max_depth == 0:
        assert (img[~miss & ~em] == 0).all()
    else:
        assert img[~miss & ~em].mean() > 0.01


@pytest.mark.parametrize("free_paths", [3, 0])
def test_pt_samples_in_flight_follow_free_memory(monkeypatch, free_paths):
    """With little free device memory (forced through the test hook) the
    path tracer keeps fewer samples in flight (3 of 8, or 1) and renders the
    oracle's image all the same."""
    s = P.Scene.cornell(3, 20000)
    t = s.build(1)
    w, h, spp, depth, rr = 64, 48, 8, 6, 2
    cam = O.camera([0, 1, 3.8], [0, 0, -1], 40, w, h)
    bg = np.asarray(s.info()["background"], np.float32)
    monkeypatch.setenv("PRAY_TEST_FREE_BYTES", str(2 * free_paths * w * h * 100))
    c = P.Context(0)  # a fresh workspace: nothing held to count as free
    try:
        sid = c.upload(t)
        say(f"pt with free memory for {free_paths} samples in flight")
        hdr, rgb = c.render_pt(sid, cam, w, h, spp, depth, rr, 77, bg, 1e-9)
        queues = c.last_render_queues()
    finally:
        c.close()
    oh, orgb, oq = O.render_pt(t.flat(), cam, w, h, spp, depth, rr, 77, bg, 1e-9, queues=True)
    assert np.array_equal(hdr.view(np.uint32), oh.view(np.uint32)) and np.array_equal(rgb, orgb)
    assert np.array_equal(queues, oq)


def test_pt_reference_cornell_example_equal_oracle(ctx):
    s, t, f = load("pt_cornell.json")
    inf = s.info()
    w, h = 80, 60
    cam = O.camera([0, 1, 0.9], [0, 0, -1], 130, w, h)
    sid = ctx.upload(t)
    try:
        say("pt reference cornell example")
        hdr, rgb = ctx.render_pt(sid, cam, w, h, 16, inf["max_depth"], 1 << 30, 5, inf["background"], 1e-4)
    finally:
        ctx.free_scene(sid)
    oh, orgb = O.render_pt(f, cam, w, h, 16, inf["max_depth"], 1 << 30, 5, inf["background"], 1e-4)
    assert np.array_equal(hdr.view(np.uint32), oh.view(np.uint32)) and np.array_equal(rgb, orgb)


def box_room(le, inner=None, rho=0.5):
    """Closed emissive cube (inward faces), optionally with an inner
    non-emissive box of albedo rho."""
    tris, col, emit = [], [], []

    def quad(p0, u, v, c, e):
        p0, u, v = (np.array(x, np.float64) for x in (p0, u, v))
        a, b, cc, d = p0, p0 + u, p0 + u + v, p0 + v
        for tri in ((a, b, cc), (a, cc, d)):
            tris.append(np.concatenate(tri))
            col.append(c)
            emit.append(e)

    L = [le] * 3
    # inward normals of the room [-2,2]^3
    quad([-2, -2, -2], [0, 0, 4], [4, 0, 0], L, 1)
    quad([-2, 2, -2], [4, 0, 0], [0, 0, 4], L, 1)
    quad([-2, -2, -2], [4, 0, 0], [0, 4, 0], L, 1)
    quad([-2, -2, 2], [0, 4, 0], [4, 0, 0], L, 1)
    quad([-2, -2, -2], [0, 4, 0], [0, 0, 4], L, 1)
    quad([2, -2, -2], [0, 0, 4], [0, 4, 0], L, 1)
    if inner is not None:
        r = inner
        R = [rho] * 3
        quad([-r, -r, -r], [4 * r / 2, 0, 0], [0, 0, 4 * r / 2], R, 0)  # bottom (-y)
        quad([-r, r, -r], [0, 0, 2 * r], [2 * r, 0, 0], R, 0)  # top (+y)
        quad([-r, -r, -r], [0, 2 * r, 0], [2 * r, 0, 0], R, 0)  # -z
        quad([-r, -r, r], [2 * r, 0, 0], [0, 2 * r, 0], R, 0)  # +z
        quad([-r, -r, -r], [0, 0, 2 * r], [0, 2 * r, 0], R, 0)  # -x
        quad([r, -r, -r], [0, 2 * r, 0], [0, 0, 2 * r], R, 0)  # +x
    return (np.array(tris, np.float32), np.array(col, np.float32), np.array(emit, np.uint8))


def test_pt_estimator_invariants(ctx):
    """All-emissive room: every pixel is Le exactly.  Room + inner box of
    albedo rho: pixels that see the box are rho*Le, the rest Le (depth >= 1,
    no roulette)."""
    le, rho = np.float32(3.0), np.float32(0.5)
    for inner in (None, 0.5):
        abc, col, emit = box_room(le, inner, rho)
        s = P.Scene.from_triangles(abc, col, emit)
        t = s.build(1)
        sid = ctx.upload(t)
        try:
            w = h = 32
            cam = O.camera([0, 0, 1.8], [0, 0, -1], 90, w, h)
            say(f"pt invariants inner={inner}")
            hdr, _ = ctx.render_pt(sid, cam, w, h, 4, 3, 1 << 30, 1, [0, 0, 0], 1e-9)
        finally:
            ctx.free_scene(sid)
        oh, _ = O.render_pt(t.flat(), cam, w, h, 4, 3, 1 << 30, 1, np.zeros(3, np.float32), 1e-9)
        assert np.array_equal(hdr.view(np.uint32), oh.view(np.uint32))
        # Moeller-Trumbore is not watertight: a ray through the exact shared
        # edge of two triangles can miss both and escape to the (black)
        # background; such pixels aside, the closed form holds exactly.
        crack = hdr == 0
        assert crack.mean() < 0.01
        if inner is None:
            assert (hdr[~crack] == le).all()
        else:
            vals = set(np.unique(hdr[~crack]).tolist())
            assert vals <= {float(le), float(np.float32(le * rho))} and len(vals) == 2


def test_cli_drop_in_renders_reference_image(tmp_path):
    name = "head_fov30_odd.json"
    out = tmp_path / "out.bmp"
    say("cli")
    r = subprocess.run([P.CLI, os.path.join(GOLDEN, "scenes", name), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    img = O.read_bmp(str(out))
    ref = np.load(os.path.join(GOLDEN, "ref_images.npz"))[f"{name}:exact"]
    assert np.array_equal(img, ref)
    # byte-identical file to the reference CLI's output
    if os.path.exists(O.REF_CLI):
        ref_out = tmp_path / "ref.bmp"
        subprocess.run([O.REF_CLI, os.path.join(GOLDEN, "scenes", name), str(ref_out)], check=True,
                       capture_output=True)
        assert out.read_bytes() == ref_out.read_bytes()


def test_pt_progressive_resume_equals_one_call(ctx):
    """Samples split over several progressive calls (the running sum passed
    on) give the one-call image bit for bit, which equals the oracle's."""
    s = P.Scene.cornell(3, 20000)
    t = s.build(1)
    w, h, spp, depth, rr = 64, 48, 8, 6, 2
    cam = O.camera([0, 1, 3.8], [0, 0, -1], 40, w, h)
    bg = np.asarray(s.info()["background"], np.float32)
    sid = ctx.upload(t)
    try:
        say("pt progressive: one call")
        hdr, rgb = ctx.render_pt(sid, cam, w, h, spp, depth, rr, 77, bg, 1e-9)
        rays_one = ctx.last_render_rays()
        for split in ([3, 5], [1, 1, 6], [7, 1]):
            say(f"pt progressive: split {split}")
            acc, first, rays = None, 0, 0
            for n in split:
                ph, prgb, acc = ctx.render_pt_progressive(sid, cam, w, h, first, n, depth, rr, 77, bg, 1e-9, accum=acc)
                rays += ctx.last_render_rays()
                first += n
                # every intermediate image is the mean of the samples so far
                assert np.array_equal(ph.view(np.uint32), (acc / np.float32(first)).view(np.uint32))
            assert np.array_equal(ph.view(np.uint32), hdr.view(np.uint32)), split
            assert np.array_equal(prgb, rgb)
            # each call traces the camera rays once for all its samples
            assert rays == rays_one + (len(split) - 1) * w * h
        # resuming needs the running sum
        with pytest.raises(P.PrayError):
            ctx._check(ctx._lib.pray_render_pt_progressive(ctx.handle, sid, P._p(cam), w, h, 3, 1, depth, rr, 77,
                                                           P._p(bg), 1e-9, None, None, None))
    finally:
        ctx.free_scene(sid)
    oh, orgb = O.render_pt(t.flat(), cam, w, h, spp, depth, rr, 77, bg, 1e-9)
    assert np.array_equal(hdr.view(np.uint32), oh.view(np.uint32)) and np.array_equal(rgb, orgb)


def run_cli(*args):
    r = subprocess.run([P.CLI, *map(str, args)], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout + r.stderr


def test_cli_progressive_accumulation_and_hdr(ctx, tmp_path):
    """`Pray scene.json out.bmp --spp N --accum state --hdr out.pfm`: two
    resumed runs (2 + 3 samples) write the same BMP and PFM as one 5-sample
    run, and the same image as the library call; a state from another
    setting is refused."""
    scene = os.path.join(GOLDEN, "scenes", "pt_cornell.json")
    st = tmp_path / "state.pacc"
    say("cli progressive 2 + 3")
    rc, out = run_cli(scene, tmp_path / "a.bmp", "--spp", 2, "--accum", st, "--hdr", tmp_path / "a.pfm")
    assert rc == 0 and "Accumulated 2 samples" in out, out
    rc, out = run_cli(scene, tmp_path / "b.bmp", "--spp", 3, "--accum", st, "--hdr", tmp_path / "b.pfm")
    assert rc == 0 and "Resuming after 2 samples" in out and "Accumulated 5 samples" in out, out
    say("cli one run of 5")
    rc, out = run_cli(scene, tmp_path / "c.bmp", "--spp", 5, "--hdr", tmp_path / "c.pfm")
    assert rc == 0, out
    assert (tmp_path / "b.bmp").read_bytes() == (tmp_path / "c.bmp").read_bytes()
    assert (tmp_path / "b.pfm").read_bytes() == (tmp_path / "c.pfm").read_bytes()
    acc, n, fp = P.load_accum(st)
    s = P.Scene.load_json(scene)
    assert n == 5 and fp == s.fingerprint()
    hdr = P.read_pfm(tmp_path / "c.pfm")
    if s.info()["max_depth"] > 1:
        assert np.array_equal(hdr.view(np.uint32), (acc / np.float32(5)).view(np.uint32))
    # (the example's depth limit of 1 takes the diffuse estimator: the state
    # holds the samples' arriving radiance and the image its product with
    # the hit's albedo, pray.cpp:315-340; the library render below, which the
    # oracle tests pin, checks that image)
    assert np.array_equal(O.read_bmp(str(tmp_path / "c.bmp")), np.asarray(
        np.clip(hdr, 0, 1) * np.float32(255) + np.float32(0.5), np.float32).astype(np.uint8))
    # the library call with the scene's own settings renders the same image
    inf = s.info()
    sid = ctx.upload(s.build(-1))
    try:
        say("library render of the same scene")
        lh, _ = ctx.render_pt(sid, s.camera(), inf["width"], inf["height"], 5, inf["max_depth"],
                              inf["rr_depth"] if inf["rr_depth"] >= 0 else 0xFFFFFFFF, inf["seed"],
                              inf["background"], inf["det_eps"])
    finally:
        ctx.free_scene(sid)
    assert np.array_equal(lh.view(np.uint32), hdr.view(np.uint32))
    # another render setting (max depth) is refused and leaves the state as it was
    import shutil
    for ext in ("obj", "mtl"):
        shutil.copy(os.path.join(GOLDEN, "scenes", f"CornellBox-Original.{ext}"), tmp_path)
    other = tmp_path / "other.json"
    other.write_text(open(scene).read().replace('"max_depth": 1', '"max_depth": 2'))
    before = st.read_bytes()
    rc, out = run_cli(other, tmp_path / "d.bmp", "--spp", 1, "--accum", st)
    assert rc != 0 and "another scene" in out, out
    assert st.read_bytes() == before


def test_free_memory_decides_gpu_eligibility(ctx, tmp_path):
    """pray_mem_ok answers from the device's free memory (here overridden
    through the library's test hook): a scene that fits is taken, the same
    scene with 1 KB free is not, and the CLI then reports and exits with
    status 2 instead of failing inside a render."""
    import subprocess
    say("mem_ok with the device's free memory and with 1 KB")
    assert ctx.mem_ok(25_602, 40_000, 20_000, 512 * 512)
    os.environ["PRAY_TEST_FREE_BYTES"] = "1024"
    try:
        assert not ctx.mem_ok(25_602, 40_000, 20_000, 512 * 512)
    finally:
        del os.environ["PRAY_TEST_FREE_BYTES"]
    assert ctx.mem_ok(25_602, 40_000, 20_000, 512 * 512)
    scene = os.path.join(GOLDEN, "scenes", "head_fov30.json")
    say("CLI with 1 KB of free device memory")
    out = tmp_path / "small.bmp"
    r = subprocess.run([P.CLI, scene, str(out)], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, PRAY_TEST_FREE_BYTES="1024"))
    assert r.returncode == 2 and "cannot take this scene" in r.stderr and not out.exists(), r.stderr
    say("CLI with the device's free memory")
    r = subprocess.run([P.CLI, scene, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and out.exists(), r.stderr
