"""Host surface: the KD-tree builders reproduce the reference's trees
(compiled from source), the flat layout is well formed, scene I/O matches
the reference loader, and the C ABI exports what include/*.h declares."""
import os
import re

import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not available")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^PRAY_API[^;(]*?\b(\w+)\s*\(", text, re.MULTILINE)))


def test_abi_exports_every_declared_symbol():
    hip, host = P.libraries()
    for name in declared("pray_hip.h"):
        assert hasattr(hip, name), name
    for name in declared("pray_host.h"):
        assert hasattr(host, name), name
    assert len(declared("pray_hip.h")) >= 20 and len(declared("pray_host.h")) >= 20


def test_tuning_array_length_matches_the_header():
    """The Python wrapper passes PRAY_TUNING_FIELDS ints to pray_set_tuning /
    pray_get_tuning: its field list must have the header's length."""
    text = open(os.path.join(ROOT, "include", "pray_hip.h")).read()
    n = int(re.search(r"^#define PRAY_TUNING_FIELDS (\d+)", text, re.MULTILINE).group(1))
    assert len(P.TUNING_FIELDS) == n
    # every field is documented in the header's list, in order
    doc = re.findall(r"^ \*   \[(\d)\] (\w+)", text, re.MULTILINE)
    assert [(int(i), f) for i, f in doc] == list(enumerate(P.TUNING_FIELDS))


def test_abi_only_exports_c_symbols():
    import subprocess
    for lib in (P.HIP_LIB, P.HOST_LIB):
        out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
        syms = [l.split()[-1] for l in out.splitlines() if " T " in l]
        assert syms and all(s.startswith(("pray_", "ph_")) for s in syms), syms


def test_context_without_gpu_fails_cleanly():
    """On a machine without a GPU, pray_init reports NODEVICE instead of
    crashing; with a GPU it succeeds."""
    try:
        c = P.Context(0)
        c.close()
    except P.PrayError as e:
        assert e.code == -6


def test_init_failure_status_reaches_the_caller(monkeypatch):
    """A failed device setup (injected through the library's test hook)
    comes back as its status code, on machines with and without a GPU."""
    for code in (-2, -6, -3):
        monkeypatch.setenv("PRAY_TEST_INIT_STATUS", str(code))
        with pytest.raises(P.PrayError) as e:
            P.Context(0)
        assert e.value.code == code


def test_cli_without_usable_gpu_reports_and_exits(tmp_path):
    """isComputableOnGPU is a bool (path_trace.cu:557-583): without a usable
    device the drop-in CLI says so and exits with status 2 -- no exception
    escapes, no signal, no image (scene_preprocessor.cpp:42-49 would fall
    back to the CPU here)."""
    import subprocess
    scene = os.path.join(ROOT, "tests", "golden", "scenes", "head_fov30.json")
    for code in ("-6", "-2"):
        out = tmp_path / f"x{code}.bmp"
        env = dict(os.environ, PRAY_TEST_INIT_STATUS=code)
        r = subprocess.run([P.CLI, scene, str(out)], capture_output=True, text=True, env=env, timeout=120)
        assert r.returncode == 2, (r.returncode, r.stderr)
        assert "cannot take this scene" in r.stderr
        assert not out.exists()


def compare_trees(scene, kind, depth=20):
    abc, _, _ = scene.triangles()
    t = scene.build(kind, max_depth=depth)
    R = O.RefDB(abc, kind)
    mine, ref = t.export(), R.export()
    for m, r in zip(mine, ref):
        assert m.shape == r.shape
        assert np.array_equal(m.view(np.uint32), r.view(np.uint32))
    st, rs = t.stats(), R.stats()
    for k in ("nnodes", "maxdepth", "leafNodes", "emptyLeafNodes", "triangleCount"):
        assert st[k] == rs[k], k
    return t


@needs_ref
@pytest.mark.parametrize("maker", [lambda: P.Scene.spheres(1), lambda: P.Scene.terrain(96, 3),
                                   lambda: P.Scene.cornell(3, 12000), lambda: P.Scene.random(4, 6000)])
@pytest.mark.parametrize("kind", [0, 1])
def test_builders_equal_reference(maker, kind):
    compare_trees(maker(), kind)


@needs_ref
def test_sah_builder_equal_reference_on_example_obj():
    s = P.Scene.load_json(os.path.join(ROOT, "tests", "golden", "scenes", "head_fov30.json"))
    compare_trees(s, 1)
    compare_trees(s, 0)


def test_flat_layout_well_formed():
    for kind in (0, 1):
        s = P.Scene.cornell(2, 5000)
        t = s.build(kind)
        f = t.flat()
        nodes = f["nodes"].reshape(-1, 2)
        refs = f["refs"]
        n_tris = f["a"].shape[0]
        assert refs[-1] & 0x80000000
        assert ((refs & 0x7FFFFFFF) < n_tris).all()
        # every triangle is referenced by at least one leaf
        assert len(np.unique(refs & 0x7FFFFFFF)) == n_tris
        seen, stack, depth, inner_lcs = set(), [(1, 0)], 0, []
        while stack:
            i, d = stack.pop()
            assert i not in seen
            seen.add(i)
            depth = max(depth, d)
            w = int(nodes[i, 1])
            if w & 3 == 3:
                continue
            lc = w >> 2
            assert lc % 2 == 0 and lc >= 2 and lc + 1 < len(nodes)
            inner_lcs.append(lc)
            stack += [(lc, d + 1), (lc + 1, d + 1)]
        assert depth == f["max_depth"]
        # treelet blocks: 16 nodes (128 bytes) each, most of them in use
        assert len(nodes) % 16 == 0 and len(seen) > 0.6 * len(nodes)
        assert len(seen) == 2 * len(inner_lcs) + 1
        # SAH children share the split plane; median children clip inside the parent
        if kind == 1:
            lcs = np.array(inner_lcs)
            assert np.array_equal(nodes[lcs, 0], nodes[lcs + 1, 0])


def test_depth_cap_extension_keeps_leaf_sets_correct():
    s = P.Scene.terrain(128, 9)
    t20 = s.build(1, max_depth=20)
    t12 = s.build(1, max_depth=12)
    assert t12.stats()["maxdepth"] <= 12
    f = t12.flat()
    rays = O.gen_rays(4, P.scene_box_expanded(f["box"]), 3000)
    a = O.flat_trace(t20.flat(), rays, 0, 1e-12)
    b = O.flat_trace(f, rays, 0, 1e-12)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("kind", [0, 1])
def test_build_parameters_change_tree_not_hits(kind):
    """Smaller leaves and a cheaper SAH traversal give a deeper tree with
    fewer triangle tests per ray and exactly the same closest hits; explicit
    reference values reproduce the default tree."""
    s = P.Scene.terrain(128, 9)
    ref = s.build(kind)
    same = s.build(kind, max_depth=20, leaf_size=32, traversal_cost=5.0, intersection_cost=1.5)
    for x, y in zip(ref.export(), same.export()):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    deep = s.build(kind, max_depth=40, leaf_size=8, traversal_cost=2.0)
    assert deep.stats()["leafNodes"] > 2 * ref.stats()["leafNodes"]
    rays = O.gen_rays(5, P.scene_box_expanded(ref.box()), 4000)
    a = O.flat_trace(ref.flat(), rays, 0, 1e-12, counts=True)
    b = O.flat_trace(deep.flat(), rays, 0, 1e-12, counts=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert b[2][:, 1].sum() < a[2][:, 1].sum()


def test_procedural_scenes_sizes_and_determinism():
    assert P.Scene.spheres(1).num_tris == 25602
    assert P.Scene.terrain(64, 1).num_tris == 2 * 64 * 64
    c = P.Scene.cornell(1, 100_000)
    assert abs(c.num_tris - 100_000) < 3000
    a1, _, _ = P.Scene.terrain(32, 5).triangles()
    a2, _, _ = P.Scene.terrain(32, 5).triangles()
    a3, _, _ = P.Scene.terrain(32, 6).triangles()
    assert np.array_equal(a1, a2) and not np.array_equal(a1, a3)
    # terrain triangles face up (det > 0 for downward rays)
    ab, ac, nn = P.Scene.terrain(32, 5).derived()
    assert (nn[:, 1] > 0).all()
    # heights within [0, 0.1]
    assert a1[:, 1::3].min() >= 0 and a1[:, 1::3].max() <= 0.1


@needs_ref
@pytest.mark.parametrize("name", ["cube_200x150.json", "head_fov30.json", "plant_256.json", "pt_cornell.json"])
def test_scene_loader_equals_reference(name):
    path = os.path.join(ROOT, "tests", "golden", "scenes", name)
    r = O.ref_load_scene(path)
    s = P.Scene.load_json(path)
    a, c, e = s.triangles()
    assert np.array_equal(a.view(np.uint32), r["tris"].view(np.uint32))
    assert np.array_equal(c.view(np.uint32), r["color"].view(np.uint32))
    assert np.array_equal(e, r["emit"])
    inf = s.info()
    assert np.array_equal(s.lights(), r["lights"])
    assert np.array_equal(inf["background"], r["background"])
    for k in ("mode", "width", "height", "fov", "max_depth", "spp"):
        assert inf[k] == r[k], k


def number_spellings(rng, n):
    """OBJ number tokens in many spellings: fixed, exponent, upper-case E,
    explicit '+', trailing '.', long fractions, signed zeros, tiny and large
    magnitudes; and tokens the OBJ number grammar rejects (read as 0)."""
    fixed = ["-0", "-0.0", "+0", "0.", "7.", "+7.5", "1e3", "1E-3", "2.5e+2", "1.e2", "123456789012.5",
             "0.1234567890123456789", "3.14159265358979323846", "1e-38", "1e-45", "4e38", "-1.17549435e-38",
             "9999999.999999999", "1e", "1e+", ".5", "-.25", "x1", "+", "1ex", "5e-x", "12abc", "0x10"]
    out = list(fixed)
    for _ in range(n):
        v = float(rng.normal() * 10.0 ** rng.uniform(-9, 9))
        f = rng.integers(0, 7)
        out.append([f"{v:.17g}", f"{v:.9f}", f"{v:e}", f"{v:.3E}", ("+" if v >= 0 else "") + f"{v:.6f}",
                    f"{v:.0f}.", f"{v:.12e}"][f])
    # Exact decimal spellings of the midpoints between adjacent floats: the
    # float a parser lands on then depends on the rounding of every step of
    # its double arithmetic, so these tell apart parsers that agree elsewhere.
    f32 = (rng.normal(size=n) * 10.0 ** rng.uniform(-6, 6, n)).astype(np.float32)
    up = np.nextafter(f32, np.float32(np.inf))
    mids = (f32.astype(np.float64) + up.astype(np.float64)) / 2
    for k, m in enumerate(mids):
        digits = int(rng.integers(8, 30))
        out.append(f"{m:.{digits}g}" if k % 2 else f"{m:.{digits}e}")
    return out


@needs_ref
def test_obj_numbers_equal_reference_loader(tmp_path):
    """Vertex coordinates and material colours written in every number
    spelling load to the same float bits as in the reference's loader
    (compiled from its sources), rejected spellings included."""
    rng = np.random.default_rng(11)
    toks = number_spellings(rng, 3000)
    toks += ["0"] * (-len(toks) % 9)
    lines = ["mtllib t.mtl", "usemtl a"]
    nv = len(toks) // 3
    for k in range(nv):
        lines.append("v " + " ".join(toks[3 * k: 3 * k + 3]))
    for k in range(nv // 3):
        lines.append(f"f {3 * k + 1} {3 * k + 2} {3 * k + 3}")
    (tmp_path / "t.obj").write_text("\n".join(lines) + "\n")
    kd = number_spellings(np.random.default_rng(12), 0)[:18]
    (tmp_path / "t.mtl").write_text("newmtl a\nKd " + " ".join(kd[:3]) + "\nKe " + " ".join(kd[3:6]) + "\n")
    js = {"obj_file": "t.obj", "resolution_x": 8, "resolution_y": 8, "camera_position": [0, 0, -5],
          "camera_look": [0, 0, 1], "fov": 60, "lights": [], "background": [0, 0, 0]}
    import json
    (tmp_path / "t.json").write_text(json.dumps(js))
    r = O.ref_load_scene(str(tmp_path / "t.json"))
    s = P.Scene.load_json(str(tmp_path / "t.json"))
    a, c, e = s.triangles()
    assert len(a) == nv // 3
    assert np.array_equal(a.view(np.uint32), r["tris"].view(np.uint32))
    assert np.array_equal(c.view(np.uint32), r["color"].view(np.uint32))
    assert np.array_equal(e, r["emit"])


def test_scene_json_round_trip(tmp_path):
    s = P.Scene.cornell(4, 2000)
    path = str(tmp_path / "c.json")
    s.write_json(path)
    r = P.Scene.load_json(path)
    a, ca, ea = s.triangles()
    b, cb, eb = r.triangles()
    assert np.array_equal(a, b) and np.array_equal(ca, cb) and np.array_equal(ea, eb)
    ia, ib = s.info(), r.info()
    for k in ("mode", "width", "height", "fov", "max_depth", "spp"):
        assert ia[k] == ib[k]
    assert np.array_equal(s.camera(), r.camera())


def test_scene_load_errors_are_reported(tmp_path):
    with pytest.raises(ValueError):
        P.Scene.load_json(str(tmp_path / "missing.json"))
    bad = tmp_path / "bad.json"
    bad.write_text("{ not json")
    with pytest.raises(ValueError):
        P.Scene.load_json(str(bad))


def test_axis_parallel_rays_keep_pruning():
    """A direction with zero components gives infinite plane distances; the
    widened interval must stay infinite (not NaN), or the ray loses all
    pruning along that axis and walks the whole tree.  Regression: a
    first-bounce path-tracing ray leaving the back wall along +z visited
    every node of the Cornell tree."""
    s = P.Scene.cornell(1, 100_000)
    f = s.build(1, max_depth=40, leaf_size=8, traversal_cost=2.0).flat()
    n_nodes = len(f["nodes"]) // 2
    rays = np.array([[0.3668138, 1.8581737, -1.0000007, 0.0, 0.0, 1.0],
                     [0.2, 0.7, 0.3, 0.0, 1.0, 0.0],
                     [0.2, 0.7, 0.3, -1.0, 0.0, 0.0],
                     [0.0, 1.0, 0.0, 0.0, -0.0, -1.0]], np.float32)
    idx, t, cnt = O.flat_trace(f, rays, 0, 1e-9, counts=True)
    bf_idx, bf_t = O.bf_closest(f["a"], f["ab"], f["ac"], rays, 1e-9)
    assert np.array_equal(idx, bf_idx) and np.array_equal(t.view(np.uint32), bf_t.view(np.uint32))
    assert cnt[:, 0].max() < 0.02 * n_nodes, cnt


def test_flat_tree_file_round_trip(tmp_path):
    """A saved flat tree loads back bit for bit and traces identically; a
    damaged or truncated file is refused."""
    s = P.Scene.cornell(5, 8000)
    t = s.build(1, max_depth=40, leaf_size=8, traversal_cost=2.0)
    path = tmp_path / "c.pflat"
    t.save(path)
    u = P.Tree.load(path)
    a, b = t.flat(), u.flat()
    for k in ("nodes", "refs", "a", "ab", "ac", "normal", "color", "emit", "box"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
    assert a["max_depth"] == b["max_depth"]
    st, su = t.stats(), u.stats()
    for k in ("kind", "maxdepth", "leafNodes", "emptyLeafNodes", "triangleCount", "flat_nodes", "flat_refs"):
        assert st[k] == su[k], k
    rays = O.gen_rays(8, P.scene_box_expanded(b["box"]), 2000)
    assert np.array_equal(O.flat_trace(a, rays, 0, 1e-9)[0], O.flat_trace(b, rays, 0, 1e-9)[0])
    raw = bytearray(path.read_bytes())
    raw[200] ^= 0x40
    bad = tmp_path / "bad.pflat"
    bad.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        P.Tree.load(bad)
    short = tmp_path / "short.pflat"
    short.write_bytes(path.read_bytes()[:-9])
    with pytest.raises(ValueError, match="length"):
        P.Tree.load(short)
    with pytest.raises(ValueError):
        P.Tree.load(tmp_path / "missing.pflat")


def parse_pfm(raw):
    """Independent PFM reader: header lines, then rows bottom first in the
    byte order given by the sign of the scale."""
    parts = raw.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = map(int, parts[1].split())
    scale = float(parts[2])
    data = np.frombuffer(parts[3], dtype="<f4" if scale < 0 else ">f4")
    assert data.size == w * h * 3
    return data.reshape(h, w, 3)[::-1].astype(np.float32)


def test_pfm_files(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.normal(size=(5, 7, 3)).astype(np.float32)
    img[0, 0] = [np.inf, -0.0, 1e-40]
    path = tmp_path / "a.pfm"
    P.write_pfm(path, img)
    raw = path.read_bytes()
    assert raw.startswith(b"PF\n7 5\n-1.0\n") and len(raw) == 12 + 5 * 7 * 12
    assert np.array_equal(parse_pfm(raw).view(np.uint32), img.view(np.uint32))
    assert np.array_equal(P.read_pfm(path).view(np.uint32), img.view(np.uint32))
    # big-endian files (positive scale) read too
    big = tmp_path / "big.pfm"
    big.write_bytes(b"PF\n7 5\n1.0\n" + img[::-1].astype(">f4").tobytes())
    assert np.array_equal(P.read_pfm(big).view(np.uint32), img.view(np.uint32))
    short = tmp_path / "short.pfm"
    short.write_bytes(raw[:-4])
    with pytest.raises(ValueError, match="truncated"):
        P.read_pfm(short)
    gray = tmp_path / "gray.pfm"
    gray.write_bytes(b"Pf\n7 5\n-1.0\n" + bytes(140))
    with pytest.raises(ValueError, match="not an RGB PFM"):
        P.read_pfm(gray)


def test_accumulation_files_and_fingerprint(tmp_path):
    s = P.Scene.load_json(os.path.join(ROOT, "tests", "golden", "scenes", "pt_cornell.json"))
    fp = s.fingerprint()
    # the sample count is not part of the fingerprint; every other setting is
    s.set_render(spp=17)
    assert s.fingerprint() == fp
    s.set_render(max_depth=3)
    assert s.fingerprint() != fp
    s2 = P.Scene.load_json(os.path.join(ROOT, "tests", "golden", "scenes", "pt_cornell.json"))
    s2.set_render(width=801)
    assert s2.fingerprint() != fp
    assert P.Scene.cornell(1, 2000).fingerprint() != P.Scene.cornell(2, 2000).fingerprint()
    acc = np.random.default_rng(4).uniform(0, 9, (6, 4, 3)).astype(np.float32)
    path = tmp_path / "s.pacc"
    P.save_accum(path, acc, 12, fp)
    got, n, gfp = P.load_accum(path)
    assert np.array_equal(got, acc) and n == 12 and gfp == fp
    raw = path.read_bytes()
    assert len(raw) == 32 + acc.nbytes + 8 and raw[:8] == b"PRAYACC1"
    assert not (tmp_path / "s.pacc.tmp").exists()
    bad = bytearray(raw)
    bad[40] ^= 1
    (tmp_path / "bad.pacc").write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="checksum"):
        P.load_accum(tmp_path / "bad.pacc")
    (tmp_path / "short.pacc").write_bytes(raw[:-1])
    with pytest.raises(ValueError, match="truncated"):
        P.load_accum(tmp_path / "short.pacc")
    (tmp_path / "long.pacc").write_bytes(raw + b"x")
    with pytest.raises(ValueError, match="trailing"):
        P.load_accum(tmp_path / "long.pacc")
    with pytest.raises(ValueError):
        P.load_accum(tmp_path / "missing.pacc")


def test_cli_rejects_bad_options(tmp_path):
    import subprocess
    scene = os.path.join(ROOT, "tests", "golden", "scenes", "head_fov30.json")
    for extra in (["--spp", "0"], ["--bogus", "1"], ["--hdr"]):
        r = subprocess.run([P.CLI, scene, str(tmp_path / "x.bmp"), *extra], capture_output=True, text=True)
        assert r.returncode != 0 and "usage" in r.stdout, extra
    # ray-tracing scenes have no running sum to keep
    r = subprocess.run([P.CLI, scene, str(tmp_path / "x.bmp"), "--hdr", str(tmp_path / "x.pfm")], capture_output=True,
                       text=True)
    assert r.returncode != 0 and "path-tracing scenes" in r.stderr
