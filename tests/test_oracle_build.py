"""The oracle's twin of the GPU tree build (a bounding interval hierarchy
over Morton-ordered triangles), checked on the CPU against what the
construction fixes: codes recomputed independently in numpy, a stable sort,
radix-tree splits at the highest differing code bit, child clip planes that
bound their subtrees, every triangle in exactly one leaf, and closest hits
equal to brute force."""
import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O


def spread3(v):
    r = np.zeros_like(v)
    for b in range(10):
        r |= ((v >> b) & 1) << (3 * b)
    return r


def morton_codes(abc):
    """Independent restatement in numpy float32 arithmetic."""
    t = abc.reshape(-1, 3, 3)
    lo, hi = t.min(1), t.max(1)
    blo, bhi = lo.min(0), hi.max(0)
    ext = (bhi - blo).astype(np.float32)
    inv = np.where(ext > 0, np.float32(1) / np.where(ext > 0, ext, 1), 0).astype(np.float32)
    c = ((lo + hi) * np.float32(0.5)).astype(np.float32)
    u = (((c - blo).astype(np.float32) * inv).astype(np.float32) * np.float32(1024)).astype(np.float32)
    q = np.clip(u, 0, 1023).astype(np.uint32)
    return (spread3(q[:, 0]) << 2) | (spread3(q[:, 1]) << 1) | spread3(q[:, 2]), lo, hi


def walk(flat):
    """(leaves as (first, count), inner entries as (pair node index, axis,
    child entries)) reached from the root."""
    nodes = flat["nodes"].reshape(-1, 2)
    refs = flat["refs"]
    leaves, inner, depth = [], [], 0
    stack = [(int(nodes[1, 1]), 0)]
    while stack:
        w, d = stack.pop()
        depth = max(depth, d)
        if w & 3 == 3:
            first = w >> 2
            k = first
            while not refs[k] & 0x80000000:
                k += 1
            leaves.append((first, k - first + 1))
            continue
        lc, axis = w >> 2, w & 3
        assert lc % 2 == 0 and lc + 1 < len(nodes)
        inner.append((lc, axis))
        stack += [(int(nodes[lc, 1]), d + 1), (int(nodes[lc + 1, 1]), d + 1)]
    return leaves, inner, depth


def subtree_tris(flat, w):
    nodes = flat["nodes"].reshape(-1, 2)
    refs = flat["refs"]
    out, stack = [], [w]
    while stack:
        w = stack.pop()
        if w & 3 == 3:
            k = w >> 2
            while True:
                out.append(int(refs[k] & 0x7FFFFFFF))
                if refs[k] & 0x80000000:
                    break
                k += 1
        else:
            lc = w >> 2
            stack += [int(nodes[lc, 1]), int(nodes[lc + 1, 1])]
    return np.array(out, np.int64)


def duplicates_scene(n):
    """n copies of one triangle: every code is equal, so every split falls
    on the sorted positions."""
    tri = np.array([0.1, 0.2, 0.3, 0.5, 0.2, 0.3, 0.1, 0.6, 0.35], np.float32)
    return np.tile(tri, (n, 1))


def flat_scene(n):
    """n random triangles in the plane y = 0.25: the scene box has no extent
    in y, so no node has a plane to bin on that axis."""
    rng = np.random.default_rng(7)
    centre = rng.uniform(0.0, 1.0, (n, 1, 2))
    xz = (centre + rng.uniform(-0.01, 0.01, (n, 3, 2))).astype(np.float32)
    out = np.empty((n, 3, 3), np.float32)
    out[:, :, 0], out[:, :, 1], out[:, :, 2] = xz[:, :, 0], np.float32(0.25), xz[:, :, 1]
    return out.reshape(n, 9)


SCENES = {
    "spheres": lambda: P.Scene.spheres(1).triangles()[0],
    "flat": lambda: flat_scene(3000),
    "terrain96": lambda: P.Scene.terrain(96, 3).triangles()[0],
    "cornell": lambda: P.Scene.cornell(3, 12000).triangles()[0],
    "random": lambda: P.Scene.random(4, 6000).triangles()[0],
    "duplicates": lambda: duplicates_scene(1000),
    "one": lambda: duplicates_scene(1),
    "two": lambda: P.Scene.random(5, 2000).triangles()[0][:2],
    "nine": lambda: P.Scene.random(6, 2000).triangles()[0][:9],
}


@pytest.mark.parametrize("name", sorted(SCENES))
@pytest.mark.parametrize("leaf_size", [1, 4, 8])
def test_lbvh_structure(name, leaf_size):
    abc = np.ascontiguousarray(SCENES[name](), np.float32).reshape(-1, 9)
    n = abc.shape[0]
    f = O.lbvh_build(abc, leaf_size)
    codes, lo, hi = morton_codes(abc)
    refs = f["refs"]
    order = refs & 0x7FFFFFFF
    # the leaf order is the stable sort of the codes
    assert np.array_equal(order, np.argsort(codes, kind="stable"))
    assert np.array_equal(f["box"], np.concatenate([lo.min(0), hi.max(0)]))
    leaves, inner, depth = walk(f)
    assert depth == f["max_depth"]
    # every sorted position in exactly one leaf, leaves at most leaf_size
    cover = np.zeros(n, np.int64)
    for first, cnt in leaves:
        cover[first:first + cnt] += 1
        assert 1 <= cnt and (cnt <= leaf_size or cnt == 1)
    assert (cover == 1).all()
    nodes = f["nodes"].reshape(-1, 2)
    skey = codes[order]
    for lc, axis in inner:
        left, right = subtree_tris(f, int(nodes[lc, 1])), subtree_tris(f, int(nodes[lc + 1, 1]))
        lb = nodes[lc, 0:1].view(np.float32)[0]
        rb = nodes[lc + 1, 0:1].view(np.float32)[0]
        # clip planes are the subtrees' exact bounds along the node's axis
        assert hi[left, axis].max() == lb and lo[right, axis].min() == rb
        # radix split: left keys all below right keys in sorted order
        lk, rk = codes[left], codes[right]
        assert lk.max() <= rk.min()
        # the split axis is that of the highest differing bit (if codes differ)
        both = np.concatenate([lk, rk])
        x = int(both.min()) ^ int(both.max())
        if x:
            bit = x.bit_length() - 1
            assert axis == {2: 0, 1: 1, 0: 2}[bit % 3]
            assert ((lk >> bit) & 1 == 0).all() and ((rk >> bit) & 1 == 1).all()
    assert skey.size == n


@pytest.mark.parametrize("name", ["spheres", "terrain96", "cornell", "random", "duplicates"])
def test_lbvh_closest_hits_equal_brute_force(name):
    abc = np.ascontiguousarray(SCENES[name](), np.float32).reshape(-1, 9)
    f = O.lbvh_build(abc, 8)
    box = f["box"]
    rays = O.gen_rays(17, P.scene_box_expanded(box), 4000)
    eps = 1e-9 if name != "spheres" else 1e-4
    gi, gt = O.flat_trace(f, rays, 0, eps)
    bi, bt = O.bf_closest(f["a"], f["ab"], f["ac"], rays, eps)
    assert np.array_equal(gi, bi)
    h = bi >= 0
    assert np.array_equal(gt[h].view(np.uint32), bt[h].view(np.uint32))
    if name != "duplicates":
        assert h.mean() > 0.01


# ---------------------------------------------------------------- SAH twin --

def walk_sah(flat):
    """Leaves (first, count, cell) and inner nodes (axis, plane, cell, depth)
    reached from the root, with cells from the splits."""
    nodes = flat["nodes"].reshape(-1, 2)
    refs = flat["refs"]
    box = flat["box"].astype(np.float32)
    leaves, inner = [], []
    stack = [(int(nodes[1, 1]), box[:3].copy(), box[3:].copy(), 0)]
    while stack:
        w, lo, hi, d = stack.pop()
        if w & 3 == 3:
            first = w >> 2
            if first == 0x3FFFFFFF:
                leaves.append((None, 0, lo, hi, d))
                continue
            k = first
            while not refs[k] & 0x80000000:
                k += 1
            leaves.append((first, k - first + 1, lo, hi, d))
            continue
        lc, a = w >> 2, w & 3
        p = nodes[lc, 0:1].view(np.float32)[0]
        assert nodes[lc + 1, 0:1].view(np.float32)[0] == p  # both children clip at the split plane
        inner.append((a, p, lo, hi, d))
        llo, lhi, rlo, rhi = lo.copy(), hi.copy(), lo.copy(), hi.copy()
        lhi[a] = p
        rlo[a] = p
        stack += [(int(nodes[lc, 1]), llo, lhi, d + 1), (int(nodes[lc + 1, 1]), rlo, rhi, d + 1)]
    return leaves, inner


def root_split(abc, leaf_size, ct, ci):
    """The root's binned-SAH plane (64 bins), restated in numpy float32."""
    t = abc.reshape(-1, 3, 3)
    lo, hi = t.min(1), t.max(1)
    blo, bhi = lo.min(0), hi.max(0)
    f32 = np.float32
    d = (bhi - blo).astype(np.float32)

    def ha(x, y, z):
        return f32(f32(f32(x * y) + f32(y * z)) + f32(z * x))

    sa = ha(d[0], d[1], d[2])
    n = lo.shape[0]
    best, best_a, best_p = f32(ci * f32(n)), -1, f32(0)
    for a in range(3):
        ext = f32(bhi[a] - blo[a])
        if not ext > 0:
            continue
        scale = f32(f32(64) / ext)
        bl = np.clip(np.floor(((lo[:, a] - blo[a]).astype(np.float32) * scale).astype(np.float32)), 0, 63).astype(int)
        bh = np.clip(np.floor(((hi[:, a] - blo[a]).astype(np.float32) * scale).astype(np.float32)), 0, 63).astype(int)
        starts, ends = np.bincount(bl, minlength=64), np.bincount(bh, minlength=64)
        for b in range(1, 64):
            nl, nr = int(starts[:b].sum()), n - int(ends[:b].sum())
            p = f32(blo[a] + f32(ext * f32(f32(b) * f32(1 / 64))))
            l, r = d.copy(), d.copy()
            l[a], r[a] = f32(p - blo[a]), f32(bhi[a] - p)
            pl, pr = f32(ha(*l) / sa), f32(ha(*r) / sa)
            cost = f32(f32(ct) + f32(f32(ci) * f32(f32(pl * f32(nl)) + f32(pr * f32(nr)))))
            if (nl == 0 or nr == 0) and not (pl == 1 or pr == 1):
                cost = f32(cost * f32(0.8))
            if cost < best:
                best, best_a, best_p = cost, a, p
    return best_a, best_p


@pytest.mark.parametrize("name", ["spheres", "terrain96", "cornell", "random", "duplicates", "one", "two", "nine"])
def test_sah_twin_structure(name):
    abc = np.ascontiguousarray(SCENES[name](), np.float32).reshape(-1, 9)
    n = abc.shape[0]
    f = O.sah_build(abc, leaf_size=4, max_depth=30, ct=2.0, ci=1.0)
    leaves, inner = walk_sah(f)
    t = abc.reshape(-1, 3, 3)
    lo, hi = t.min(1), t.max(1)
    assert max([lv[4] for lv in leaves]) == f["max_depth"] <= 30
    seen = np.zeros(n, np.int64)
    for first, cnt, clo, chi, d in leaves:
        if first is None:
            continue
        tri = f["refs"][first:first + cnt] & 0x7FFFFFFF
        assert (np.diff(tri.astype(np.int64)) > 0).all()  # stable partitions keep triangle order
        seen[tri] += 1
        # a leaf holds only triangles whose box meets its cell
        assert (lo[tri] <= chi).all() and (hi[tri] >= clo).all()
    assert (seen >= 1).all()
    # every triangle whose box meets a leaf's cell interior is in that leaf
    for first, cnt, clo, chi, d in leaves[:200]:
        inside = ((lo < chi) & (hi > clo)).all(1)
        got = set() if first is None else set((f["refs"][first:first + cnt] & 0x7FFFFFFF).tolist())
        assert set(np.flatnonzero(inside).tolist()) <= got
    # the root's plane is the binned-SAH minimum restated in numpy
    a, p = root_split(abc, 4, 2.0, 1.0)
    if a < 0 or n <= 4:
        assert not inner
    else:
        assert inner[0][4] == 0 and inner[0][0] == a and inner[0][1] == p


@pytest.mark.parametrize("name", ["spheres", "terrain96", "cornell", "random", "duplicates"])
def test_sah_twin_closest_hits_equal_brute_force(name):
    abc = np.ascontiguousarray(SCENES[name](), np.float32).reshape(-1, 9)
    f = O.sah_build(abc)
    rays = O.gen_rays(19, P.scene_box_expanded(f["box"]), 4000)
    eps = 1e-9 if name != "spheres" else 1e-4
    gi, gt = O.flat_trace(f, rays, 0, eps)
    bi, bt = O.bf_closest(f["a"], f["ab"], f["ac"], rays, eps)
    assert np.array_equal(gi, bi)
    h = bi >= 0
    assert np.array_equal(gt[h].view(np.uint32), bt[h].view(np.uint32))


def canonical(flat):
    """Depth-first description of a flat tree, independent of its node
    order: inner nodes (axis, clip planes) and leaf triangle lists."""
    nodes = flat["nodes"].reshape(-1, 2)
    refs = flat["refs"]
    out, stack = [], [int(nodes[1, 1])]
    while stack:
        w = stack.pop()
        if w & 3 == 3:
            first = w >> 2
            if first == 0x3FFFFFFF:
                out.append(("E",))
                continue
            k, lst = first, []
            while True:
                lst.append(int(refs[k] & 0x7FFFFFFF))
                if refs[k] & 0x80000000:
                    break
                k += 1
            out.append(("L", tuple(lst)))
            continue
        lc = w >> 2
        out.append(("N", w & 3, int(nodes[lc, 0]), int(nodes[lc + 1, 0])))
        stack += [int(nodes[lc + 1, 1]), int(nodes[lc, 1])]
    return out


@pytest.mark.parametrize("name", ["spheres", "terrain96", "cornell", "random", "duplicates", "two", "nine"])
@pytest.mark.parametrize("leaf_size,depth", [(8, 40), (1, 24)])
def test_cpu_binned_builder_equals_twin(name, leaf_size, depth):
    """The product's CPU binned-SAH builder (ph_build_ex kind 2) makes the
    twin's tree node for node; only the node order (treelet layout)
    differs."""
    abc = np.ascontiguousarray(SCENES[name](), np.float32).reshape(-1, 9)
    s = P.Scene.from_triangles(abc)
    t = s.build(2, max_depth=depth, leaf_size=leaf_size, traversal_cost=2.0)
    f = t.flat()
    o = O.sah_build(abc, leaf_size, depth, 2.0, 1.0)
    assert canonical(f) == canonical(o)
    assert f["max_depth"] == o["max_depth"] and t.stats()["kind"] == 2
