"""Pins the oracle (CPU restatement of the reference) before it is trusted:
known-answer vectors, closed forms, invariants, and equality with the
reference's own sources compiled by oracle/Makefile."""
import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not available")


# ------------------------------------------------------------- Philox KATs --
# Random123 known-answer vectors for philox4x32_10 (kat_vectors in the
# Random123 distribution; Salmon et al., SC 2011).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert list(O.philox(ctr, key)) == want


def test_sincos_accuracy_and_identity():
    u = np.concatenate([np.linspace(0, 1, 100001, endpoint=False, dtype=np.float32),
                        np.array([0.0, 0.25, 0.5, 0.75, 0.125], np.float32)])
    s, c = O.sincos2pi(u)
    ang = 2 * np.pi * u.astype(np.float64)
    assert np.max(np.abs(s - np.sin(ang))) < 2e-7
    assert np.max(np.abs(c - np.cos(ang))) < 2e-7
    assert np.max(np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1)) < 5e-7
    # exact quadrant points
    s4, c4 = O.sincos2pi(np.array([0.0, 0.25, 0.5, 0.75], np.float32))
    assert list(s4) == [0.0, 1.0, -0.0, -1.0] and list(c4) == [1.0, -0.0, -1.0, 0.0]


def test_gen_rays_distribution():
    box = np.array([-1, -2, -3, 2, 4, 6], np.float32)
    r = O.gen_rays(99, box, 200_000)
    o, d = r[:, :3], r[:, 3:]
    assert (o >= box[:3]).all() and (o <= box[:3] + box[3:]).all()
    assert np.allclose(np.linalg.norm(d.astype(np.float64), axis=1), 1, atol=1e-6)
    # uniform on the sphere: each component has mean 0 and variance 1/3
    assert np.allclose(d.mean(0), 0, atol=0.01) and np.allclose(d.var(0), 1 / 3, atol=0.01)
    # different seeds give different streams, same seed the same
    assert not np.array_equal(r, O.gen_rays(100, box, 200_000))
    assert np.array_equal(r, O.gen_rays(99, box, 200_000))


# ------------------------------------------------------------------ MT ------
def tri(a, b, c):
    a, b, c = (np.array(v, np.float32) for v in (a, b, c))
    return a, b - a, c - a


def test_mt_closed_form_axis_aligned():
    # z = 2 plane triangle facing -z, ray from origin along +z: exact t = 2.
    a, ab, ac = tri([0, 0, 2], [0, 1, 2], [1, 0, 2])  # normal (ab x ac) points to -z
    h, t = O.mt(np.array([0.25, 0.25, 0, 0, 0, 1], np.float32), a, ab, ac, 1e-4)
    assert h and t == 2.0
    # same triangle, reversed winding: culled (det < eps)
    a2, ab2, ac2 = tri([0, 0, 2], [1, 0, 2], [0, 1, 2])
    h, _ = O.mt(np.array([0.25, 0.25, 0, 0, 0, 1], np.float32), a2, ab2, ac2, 1e-4)
    assert not h
    # ray behind: t negative is still reported by MT (callers require t > 0)
    h, t = O.mt(np.array([0.25, 0.25, 5, 0, 0, 1], np.float32), a, ab, ac, 1e-4)
    assert h and t == -3.0


def test_mt_edges_inclusive_and_outside():
    a, ab, ac = tri([0, 0, 1], [0, 1, 1], [1, 0, 1])
    for p, want in [((0, 0), True), ((1, 0), True), ((0, 1), True), ((0.5, 0.5), True),
                    ((0.5, 0.5000001), False), ((-1e-6, 0.5), False), ((0.3, 0.3), True)]:
        h, t = O.mt(np.array([p[0], p[1], 0, 0, 0, 1], np.float32), a, ab, ac, 1e-4)
        assert h == want, (p, h)
        if h:
            assert t == 1.0


def test_mt_det_cull_boundary():
    # det = ab . (d x ac); for d = +z and a unit right triangle det = 1 * scale^2
    s = np.float32(0.01)  # det = 1e-4 exactly representable? use comparison with eps
    a, ab, ac = tri([0, 0, 1], [0, s, 1], [s, 0, 1])
    ray = np.array([0.001, 0.001, 0, 0, 0, 1], np.float32)
    det = float(np.dot(ab, np.cross(ray[3:], ac)))
    h_lo, _ = O.mt(ray, a, ab, ac, np.float32(det))  # det < det is false -> kept
    h_hi, _ = O.mt(ray, a, ab, ac, np.nextafter(np.float32(det), np.float32(1)))
    assert h_lo and not h_hi


def test_mt_scalar_equals_avx_bitwise():
    rng = np.random.default_rng(0)
    for _ in range(20):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        a, ab, ac = v[0], v[1] - v[0], v[2] - v[0]
        rays = O.gen_rays(int(rng.integers(1 << 30)), np.array([-2, -2, -2, 4, 4, 4], np.float32), 4000)
        hit_v, t_v = O.mt_avx(rays, a, ab, ac, 1e-4)
        for k in range(0, 4000, 7):
            h, t = O.mt(rays[k], a, ab, ac, 1e-4)
            assert h == hit_v[k]
            if h:
                assert np.float32(t).view(np.uint32) == t_v[k].view(np.uint32)


# ---------------------------------------------------------------- slab -----
def test_slab_closed_forms():
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    rays = np.array([
        [0.5, 0.5, -1, 0, 0, 1],     # through the box
        [0.5, 0.5, 2, 0, 0, 1],      # behind the box: the LINE still crosses it
        [2, 0.5, 0.5, 0, 1, 0],      # parallel, outside the x slab
        [0.5, 2, 0.5, 1, 0, 0],      # parallel, outside the y slab
        [-1, -1, -1, 0.57735026, 0.57735026, 0.57735026],
        [1.5, 0.5, 0.5, 0, 0, 1],    # parallel to z, outside in x
    ], np.float32)
    s, a = O.slab(rays, box)
    assert list(s) == [True, True, False, False, True, False]
    assert list(a) == list(s)


# ----------------------------------------------- equality with reference ---
def scene_tris(n_side=24, seed=1):
    import nostrum_pray_amd as P
    s = P.Scene.terrain(n_side, seed)
    return s


@needs_ref
def test_init_tris_equal_reference():
    import nostrum_pray_amd as P
    abc, _, _ = P.Scene.spheres(3).triangles()
    mine = O.init_tris(abc)
    ref = O.ref_init_tris(abc)
    for m, r in zip(mine, ref):
        assert np.array_equal(m.view(np.uint32), r.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("pos,look,fov,w,h", [([0, 3, 8], [0, -0.35, -1], 60, 64, 48),
                                               ([2, 3, 3], [0, 0, 1], 90, 40, 30),
                                               ([0, 1, 0.9], [0, 0, -1], 130, 33, 17),
                                               ([1.5, 5, 6], [-0.1, -0.5, -1], 30, 20, 20)])
def test_camera_equals_reference(pos, look, fov, w, h):
    cam = O.camera(pos, look, fov, w, h)
    mine = O.camera_rays(cam, w, h)
    ref = O.ref_prim_rays(pos, look, fov, w, h)
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("scene_kind", ["spheres", "cornell"])
def test_bf_and_exhaustive_equal_reference(scene_kind):
    import nostrum_pray_amd as P
    s = P.Scene.spheres(1) if scene_kind == "spheres" else P.Scene.cornell(5, 6000)
    abc, _, _ = s.triangles()
    ab, ac, _ = s.derived()
    a = abc[:, :3].copy()
    t = s.build(0)
    rays = np.concatenate([s.prim_rays()[::97], O.gen_rays(3, P.scene_box_expanded(t.box()), 3000)])
    bi, bt = O.bf_closest(a, ab, ac, rays, 1e-4)
    R = O.RefDB(abc, 2)
    ri, rt = R.intersect(rays)
    assert np.array_equal(bi, ri)
    assert np.array_equal(bt[bi >= 0], rt[ri >= 0])
    KD = O.RefDB(abc, 0)
    ki, kt = KD.intersect(rays)
    ei, et = O.ref_closest(t.export(), a, ab, ac, rays, 1e-4)
    assert np.array_equal(ei, ki) and np.array_equal(et[ei >= 0], kt[ki >= 0])
    kai, kat, _ = KD.intersect_avx(rays, 2)
    oi, ot, _ = O.ref_closest_avx(t.export(), a, ab, ac, rays, 1e-4, 2)
    assert np.array_equal(oi, kai) and np.array_equal(ot[oi >= 0], kat[kai >= 0])
    # the exhaustive traversal equals brute force except at equal-t ties,
    # where the reference's depth-first order may pick another triangle and
    # brute force (strict <, index order) picks the lowest index
    diff = np.flatnonzero(ei != bi)
    assert np.array_equal(et.view(np.uint32), bt.view(np.uint32))
    assert ((bi[diff] >= 0) & (ei[diff] >= 0) & (bi[diff] < ei[diff])).all()
    assert diff.size <= rays.shape[0] // 100


def test_flat_twin_equals_bruteforce():
    import nostrum_pray_amd as P
    for kind in (0, 1):
        s = P.Scene.cornell(7, 8000)
        t = s.build(kind)
        f = t.flat()
        rays = np.concatenate([s.prim_rays()[::61], O.gen_rays(11, P.scene_box_expanded(f["box"]), 4000)])
        for eps in (1e-4, 1e-9):
            fi, ft = O.flat_trace(f, rays, 0, eps)
            bi, bt = O.bf_closest(f["a"], f["ab"], f["ac"], rays, eps)
            assert np.array_equal(fi, bi)
            assert np.array_equal(ft[fi >= 0], bt[bi >= 0])


def leaf_runs(refs, k):
    """Triangle indices of the leaf run starting at k."""
    out = []
    while True:
        r = int(refs[k])
        out.append(r & 0x7FFFFFFF)
        if r & 0x80000000:
            return out
        k += 1


def boxed_subtrees(L):
    """Per pair of a boxed layout, for each child: the triangles its subtree
    references (by walking the records), and the leaves' run lengths."""
    rec, refs = L["rec"], L["refs"]
    runs = []

    def tris(w):
        if w == 0xFFFFFFFF:
            return []
        if w & 3 == 3:
            r = leaf_runs(refs, w >> 2)
            runs.append(len(r))
            return r
        p = w >> 3
        return tris(int(rec[p, 12])) + tris(int(rec[p, 13]))

    return tris, runs


def test_boxed_layout_structure():
    """The library's boxed layout (pray_boxed_layout, host code) of a median
    tree: it references every leaf reference of the tree once (a partition of
    each leaf), its leaves hold at most 2, every child box is the union of its
    triangles' boxes with the vertices a + ab, a + ac summed exactly (float64)
    and rounded outward -- the largest float at or below the exact minimum and
    the smallest at or above the exact maximum (a median tree's children
    pairs have two clips, so nothing is cut to a cell)."""
    import nostrum_pray_amd as P
    s = P.Scene.spheres(3)
    f = s.build(0).flat()
    L = P.boxed_layout(f)
    rec, refs = L["rec"], L["refs"]
    assert np.array_equal(np.sort(refs & 0x7FFFFFFF), np.sort(f["refs"] & 0x7FFFFFFF))
    assert rec[0, 12] == 0xFFFFFFFF and L["depth"] <= f["max_depth"] + 5
    a, ab, ac = (f[k].astype(np.float64).reshape(-1, 3) for k in ("a", "ab", "ac"))
    v = np.stack([a, a + ab, a + ac], 1)
    tmin, tmax = v.min(1), v.max(1)
    tris, runs = boxed_subtrees(L)
    fb = rec[:, :12].copy().view(np.float32)
    checked = 0
    for p in range(len(rec)):
        for ch in (0, 1):
            t = tris(int(rec[p, 12 + ch]))
            lo = fb[p, [ch, 4 + ch, 8 + ch]].astype(np.float32)
            hi = fb[p, [2 + ch, 6 + ch, 10 + ch]].astype(np.float32)
            if not t:
                assert (lo == np.inf).all() and (hi == -np.inf).all()
                continue
            elo, ehi = tmin[t].min(0), tmax[t].max(0)
            assert (lo <= elo).all() and (np.nextafter(lo, np.float32(np.inf)) > elo).all(), (p, ch)
            assert (hi >= ehi).all() and (np.nextafter(hi, np.float32(-np.inf)) < ehi).all(), (p, ch)
            checked += 1
    assert checked > 1000
    assert max(runs) <= 2 and sum(runs) >= len(refs)
    # the root box is the scene's box up to that rounding
    assert np.allclose(fb[0, [1, 5, 9, 3, 7, 11]], f["box"], rtol=1e-6, atol=1e-6)


def rounded_box(tmin, tmax):
    """Float bounds of exact (float64) bounds: the largest float at or below
    each minimum and the smallest at or above each maximum."""
    lo, hi = tmin.astype(np.float32), tmax.astype(np.float32)
    lo = np.where(lo.astype(np.float64) > tmin, np.nextafter(lo, np.float32(-np.inf)), lo)
    hi = np.where(hi.astype(np.float64) < tmax, np.nextafter(hi, np.float32(np.inf)), hi)
    return lo, hi


@pytest.mark.parametrize("kind", [1, 2])
def test_boxed_layout_cut_to_kd_cells(kind):
    """The boxed layout of a SAH KD tree (equal clips: split planes): walking
    the uploaded tree from the root with each node's cell (child 0 at or
    below its pair's plane on the parent's axis, child 1 at or above it),
    every pair of the layout made for an inner node holds its children's
    boxes as the recursion over the uploaded tree gives them: a leaf's box is
    the rounded box of its triangles cut to its cell (empty -> {+inf, -inf}),
    an inner node's the union of its children's.  Cuts happen (the boxes are
    smaller than their triangles'), and never cut across a plane the wrong
    way (a child 0 box never reaches above its plane)."""
    import nostrum_pray_amd as P
    s = P.Scene.cornell(2, 6000)
    f = s.build(kind).flat()
    L = P.boxed_layout(f)
    rec = L["rec"]
    fb = rec[:, :12].copy().view(np.float32)
    nodes = np.asarray(f["nodes"], np.uint32).reshape(-1)
    refs = np.asarray(f["refs"], np.uint32)
    a, ab, ac = (f[k].astype(np.float64).reshape(-1, 3) for k in ("a", "ab", "ac"))
    v = np.stack([a, a + ab, a + ac], 1)
    tlo, thi = rounded_box(v.min(1), v.max(1))
    E = 0xFFFFFFFF
    empty = (np.full(3, np.inf, np.float32), np.full(3, -np.inf, np.float32))
    stats = {"pairs": 0, "cut": 0, "planes": 0}

    def leaf_tris(w):
        k, out = int(w) >> 2, []
        if k == 0x3FFFFFFF:
            return out
        while True:
            out.append(int(refs[k]) & 0x7FFFFFFF)
            if refs[k] & 0x80000000:
                return out
            k += 1

    def box_of(i, clo, chi, bw):
        """Expected box of uploaded node i in cell [clo, chi]; bw is its word
        in the layout, whose pairs are checked on the way."""
        w = int(nodes[2 * i + 1])
        if w & 3 == 3:
            t = leaf_tris(w)
            if not t:
                return empty
            lo = np.maximum(tlo[t].min(0), clo)
            hi = np.minimum(thi[t].max(0), chi)
            if (tlo[t].min(0) < clo).any() or (thi[t].max(0) > chi).any():
                stats["cut"] += 1
            return (lo, hi) if (lo <= hi).all() else empty
        axis, lc = w & 3, w >> 2
        p = bw >> 3
        c0lo, c0hi, c1lo, c1hi = clo.copy(), chi.copy(), clo.copy(), chi.copy()
        assert nodes[2 * lc] == nodes[2 * lc + 2]  # a SAH tree's split plane
        plane = np.float32(nodes[2 * lc : 2 * lc + 1].view(np.float32)[0])
        c0hi[axis] = min(c0hi[axis], plane)
        c1lo[axis] = max(c1lo[axis], plane)
        stats["planes"] += 1
        b0 = box_of(lc, c0lo, c0hi, int(rec[p, 12]))
        b1 = box_of(lc + 1, c1lo, c1hi, int(rec[p, 13]))
        for ch, (lo, hi) in ((0, b0), (1, b1)):
            got_lo = fb[p, [ch, 4 + ch, 8 + ch]]
            got_hi = fb[p, [2 + ch, 6 + ch, 10 + ch]]
            assert np.array_equal(got_lo, lo) and np.array_equal(got_hi, hi), (i, ch, got_lo, lo, got_hi, hi)
        if np.isfinite(b0[1][axis]):
            assert b0[1][axis] <= plane
        if np.isfinite(b1[0][axis]):
            assert b1[0][axis] >= plane
        stats["pairs"] += 1
        return np.minimum(b0[0], b1[0]), np.maximum(b0[1], b1[1])

    inf = np.full(3, np.inf, np.float32)
    root = int(rec[0, 13])
    assert root != E and root & 3 != 3
    box_of(1, -inf, inf, root)
    assert stats["pairs"] > 200 and stats["cut"] > 100, stats


def test_boxed_layout_4wide_structure():
    """The 4-wide form of the boxed layout: every record's live slots are
    exactly the binary layout's children and grandchildren two levels down
    (a leaf child, or both children of an inner child), with their boxes and
    leaf words, so that both forms reach the same leaves; unused slots are
    empty; the stack bound is three entries per record level."""
    import nostrum_pray_amd as P
    f = P.Scene.spheres(3).build(0).flat()
    L2, L4 = P.boxed_layout(f, 2), P.boxed_layout(f, 4)
    assert np.array_equal(L2["refs"], L4["refs"])
    r2, r4 = L2["rec"], L4["rec"]
    E = 0xFFFFFFFF

    def slots2(w):  # the binary layout's slots of inner binary word w
        out = []
        p = w >> 3
        for ch in (0, 1):
            cw = int(r2[p, 12 + ch])
            box = r2[p, [ch, 2 + ch, 4 + ch, 6 + ch, 8 + ch, 10 + ch]]
            if cw != E and cw & 3 != 3:
                q = cw >> 3
                out += [(int(r2[q, 12 + c]), r2[q, [c, 2 + c, 4 + c, 6 + c, 8 + c, 10 + c]]) for c in (0, 1)]
            else:
                out.append((cw, box))
        return out

    levels, leaves2, leaves4 = [0], [], []

    def walk(w4, w2, d):
        levels[0] = max(levels[0], d)
        q = w4 >> 3
        s = slots2(w2)
        assert 2 <= len(s) <= 4
        for j in range(4):
            w = int(r4[q, 24 + j])
            box4 = r4[q, [j, 4 + j, 8 + j, 12 + j, 16 + j, 20 + j]]
            if j >= len(s):
                assert w == E and (box4[[0, 2, 4]] == 0x7F800000).all() and (box4[[1, 3, 5]] == 0xFF800000).all()
                continue
            sw, sbox = s[j]
            assert np.array_equal(box4, sbox)
            if sw == E or sw & 3 == 3:
                assert w == sw
                if sw != E:
                    leaves4.append(w)
            else:
                walk(w, sw, d + 1)

    root2 = int(r2[0, 13])
    walk(L4["root"], root2, 0)
    assert L4["depth"] == 3 * (levels[0] + 1)

    def leaves_of2(w):
        if w == E:
            return
        if w & 3 == 3:
            leaves2.append(w)
            return
        for c in (0, 1):
            leaves_of2(int(r2[w >> 3, 12 + c]))

    leaves_of2(root2)
    assert sorted(leaves2) == sorted(leaves4) and len(leaves4) > 1000


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("which", ["spheres", "random", "cornell", "cornell-sah", "terrain-sah"])
def test_boxed_twin_equals_bruteforce(which, width):
    """The boxed node-step twin over the library's boxed layout of median
    trees and of SAH KD trees (boxes cut to their cells) gives the
    brute-force closest hit on every ray, camera and incoherent, and any-hit
    answers that threshold it, with fewer triangle tests than the plane
    form."""
    import nostrum_pray_amd as P
    s = {"spheres": lambda: P.Scene.spheres(1), "random": lambda: P.Scene.random(3, 5000),
         "cornell": lambda: P.Scene.cornell(7, 8000), "cornell-sah": lambda: P.Scene.cornell(7, 8000),
         "terrain-sah": lambda: P.Scene.terrain(64, 9)}[which]()
    f = (s.build(2, max_depth=40, leaf_size=8, traversal_cost=2.0) if which.endswith("-sah") else s.build(0)).flat()
    L = P.boxed_layout(f, width)
    rays = np.concatenate([s.prim_rays()[::41], O.gen_rays(13, P.scene_box_expanded(f["box"]), 6000)])
    for eps in (1e-4, 1e-9):
        fi, ft, fc = O.flat_trace(f, rays, 0, eps, counts=True, boxed=L)
        bi, bt = O.bf_closest(f["a"], f["ab"], f["ac"], rays, eps)
        assert np.array_equal(fi, bi)
        assert np.array_equal(ft[fi >= 0], bt[bi >= 0])
        _, _, pc = O.flat_trace(f, rays, 0, eps, counts=True)
        assert fc[:, 1].sum() < pc[:, 1].sum()
    hit = fi >= 0
    lim = np.ones_like(ft)
    lim[hit] = ft[hit] * ft[hit]
    ai, _ = O.flat_trace(f, rays, 1, 1e-9, limit2=lim, boxed=L)
    ii, _ = O.flat_trace(f, rays, 2, 1e-9, limit2=lim, boxed=L)
    assert not (ai >= 0).any()
    assert np.array_equal(ii >= 0, hit)
    lim2 = np.where(hit, lim * np.float32(4.0), np.float32(1.0)).astype(np.float32)
    ai2, _ = O.flat_trace(f, rays, 1, 1e-9, limit2=lim2, boxed=L)
    assert np.array_equal(ai2 >= 0, hit)


def test_flat_twin_any_hit_is_threshold_of_closest():
    import nostrum_pray_amd as P
    s = P.Scene.spheres(2)
    t = s.build(1)
    f = t.flat()
    rays = O.gen_rays(5, P.scene_box_expanded(f["box"]), 20000)
    ci, ct = O.flat_trace(f, rays, 0, 1e-4)
    lim = np.ones_like(ct)
    hit = ci >= 0
    lim[hit] = ct[hit] * ct[hit]  # misses keep 1 (their t is FLT_MAX)
    ai, _ = O.flat_trace(f, rays, 1, 1e-4, limit2=lim)
    ii, _ = O.flat_trace(f, rays, 2, 1e-4, limit2=lim)
    assert not (ai >= 0).any()  # strict: t*t < t*t never holds
    assert np.array_equal(ii >= 0, ci >= 0)  # inclusive: exactly the hit rays


def test_pt_oracle_closed_forms():
    """Estimator invariants of the restated path tracer (pray.cpp:342-378):
    inside an all-emissive room every pixel is Le; with an inner box of
    albedo rho, pixels are rho*Le (box) or Le (wall); with depth 0 the box
    pixels are 0 (depth limit) and with roulette the mean stays unbiased."""
    import nostrum_pray_amd as P
    from test_gpu_render import box_room
    le, rho = np.float32(3.0), np.float32(0.5)
    w = h = 24
    cam = O.camera([0, 0, 1.8], [0, 0, -1], 90, w, h)
    abc, col, emit = box_room(le)
    f = P.Scene.from_triangles(abc, col, emit).build(1).flat()
    hdr, rgb = O.render_pt(f, cam, w, h, 3, 2, 1 << 30, 1, [0, 0, 0], 1e-9)
    assert (hdr == le).all() and (rgb == 255).all()
    abc, col, emit = box_room(le, 0.5, rho)
    f = P.Scene.from_triangles(abc, col, emit).build(1).flat()
    hdr, _ = O.render_pt(f, cam, w, h, 3, 2, 1 << 30, 1, [0, 0, 0], 1e-9)
    box = hdr[..., 0] == np.float32(le * rho)
    assert box.any() and (box | (hdr[..., 0] == le)).all()
    hdr0, _ = O.render_pt(f, cam, w, h, 3, 0, 1 << 30, 1, [0, 0, 0], 1e-9)
    assert (hdr0[box] == 0).all() and (hdr0[~box] == le).all()
    # Russian roulette from bounce 0 with q = rho: survivors carry 1/q, so the
    # mean over many samples approaches rho * Le / rho * rho = rho * Le
    hdr_rr, _ = O.render_pt(f, cam, w, h, 256, 2, 0, 1, [0, 0, 0], 1e-9)
    assert abs(hdr_rr[box].mean() - le * rho) < 0.05 * le * rho
    # queue sizes: every camera path is traced at bounce 0; exactly the box
    # pixels' paths continue to bounce 1 (the walls emit), and all of those
    # reach bounce 2 (the box sees only emitting walls, no roulette)
    _, _, q = O.render_pt(f, cam, w, h, 3, 2, 1 << 30, 1, [0, 0, 0], 1e-9, queues=True)
    assert q[0] == w * h * 3 and q[1] == 3 * box.sum() and q[2] == 0
    _, _, q0 = O.render_pt(f, cam, w, h, 3, 0, 1 << 30, 1, [0, 0, 0], 1e-9, queues=True)
    assert list(q0) == [w * h * 3]


def test_pt_oracle_diffuse_estimator_closed_forms():
    """The DIFFUSE estimator the reference takes at a depth limit of 1
    (pray.cpp:315-340): in the emissive room with an inner box every box
    pixel is (Le + Le + Le) * rho / 3 in that order of float operations --
    which for these values differs from NORMAL's sum of rho * Le / 3 -- and
    every wall pixel is Le itself; no roulette even from bounce 0; every
    camera path is counted at bounce 0 and the box pixels' at bounce 1."""
    import nostrum_pray_amd as P
    from test_gpu_render import box_room
    f32 = np.float32
    le, rho, n = f32(0.9), f32(0.61), 3
    w = h = 24
    cam = O.camera([0, 0, 1.8], [0, 0, -1], 90, w, h)
    abc, col, emit = box_room(le, 0.5, rho)
    f = P.Scene.from_triangles(abc, col, emit).build(1).flat()
    hdr, _, q = O.render_pt(f, cam, w, h, n, 1, 0, 1, [0, 0, 0], 1e-9, queues=True)
    s = f32(0)
    for _ in range(n):
        s = f32(s + le)
    diffuse = f32(f32(s * rho) / f32(n))
    t = f32(0)
    for _ in range(n):
        t = f32(t + f32(rho * le))
    normal = f32(t / f32(n))
    assert diffuse != normal  # the check tells the two orders apart
    box = hdr[..., 0] != le
    assert box.any() and (hdr[box] == diffuse).all() and (hdr[~box] == le).all()
    assert list(q) == [w * h * n, n * int(box.sum())]


def test_sample_hemisphere_properties():
    rng = np.random.default_rng(1)
    n = rng.normal(size=(50000, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    u1 = rng.random(50000, dtype=np.float32)
    u2 = rng.random(50000, dtype=np.float32)
    d = O.sample_hemisphere(n, u1, u2)
    assert np.allclose(np.linalg.norm(d.astype(np.float64), axis=1), 1, atol=1e-6)
    cos = (d.astype(np.float64) * n).sum(1)
    assert (cos >= -1e-6).all()
    # cosine-weighted: cos(theta) = sqrt(1 - u1) (up to rounding), so
    # cos^2 is uniform with mean 1/2
    assert np.allclose(cos, np.sqrt(1 - u1.astype(np.float64)), atol=2e-5)
    assert abs((cos ** 2).mean() - 0.5) < 0.01
