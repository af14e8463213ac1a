"""ORACLE -- test infrastructure only.

ctypes wrappers of the CPU restatement of the reference (``liboracle.so``,
built from ``oracle.cpp``) and of the reference's own sources compiled by
``oracle/Makefile`` into ``oracle/_ref/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
this package, as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB = os.path.join(REF_DIR, "libpray_ref.so")
REF_LIB_FMA = os.path.join(REF_DIR, "libpray_ref_fma.so")
REF_CLI = os.path.join(REF_DIR, "Pray")
REF_CLI_FMA = os.path.join(REF_DIR, "Pray_fma")

_vp, _u32, _u64, _f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float
_lib = None
_refs = {}


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} not built (make -C oracle oracle)")
        L = C.CDLL(ORACLE_LIB)
        sig = {
            "orc_philox": (None, [_vp, _u32, _u32, _vp]),
            "orc_sincos2pi": (None, [_vp, _u64, _vp, _vp]),
            "orc_sample_hemisphere": (None, [_vp, _vp, _vp, _u64, _vp]),
            "orc_gen_rays": (None, [_u64, _vp, _u64, _vp]),
            "orc_init_tris": (None, [_vp, _u64, _vp, _vp, _vp]),
            "orc_mt": (C.c_int, [_vp, _vp, _vp, _vp, _f32, _vp]),
            "orc_mt_avx": (None, [_vp, _u64, _vp, _vp, _vp, _f32, _vp, _vp]),
            "orc_slab": (None, [_vp, _u64, _vp, _vp, _vp]),
            "orc_bf_closest": (None, [_vp, _vp, _vp, _u64, _vp, _u64, _f32, _vp, _vp]),
            "orc_ref_closest": (None, [_vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _u64, _f32, _vp, _vp]),
            "orc_ref_closest_avx": (C.c_double, [_vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _u64, _f32, C.c_int, _vp,
                                                 _vp]),
            "orc_flat_trace": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _u64, C.c_int, _f32, _vp, _vp, _vp,
                                      _vp]),
            "orc_trace_boxed": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _u64, C.c_int, _f32, _vp, _vp, _vp,
                                       _vp]),
            "orc_trace_boxed4": (None, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _u64, C.c_int, _f32, _vp, _vp,
                                        _vp, _vp]),
            "orc_camera": (None, [_vp, _vp, _u64, _u64, _u64, _vp]),
            "orc_camera_rays": (None, [_vp, _u64, _u64, _vp]),
            "orc_mlp_encode": (None, [_vp, _u64, _vp]),
            "orc_mlp_forward": (None, [_vp, _vp, _u64, C.c_int, _vp]),
            "orc_mlp_grad": (None, [_vp, _vp, _vp, _u64, C.c_int, _vp, _vp, _vp]),
            "orc_adam": (None, [_vp, _vp, _vp, _vp, _u64, _u32, _f32]),
            "orc_lbvh_build": (_u32, [_vp, _u64, _u32, _vp, _vp, _vp]),
            "orc_sah_build": (_u64, [_vp, _u64, _u32, _u32, _f32, _f32, _vp, _u64, _vp, _u64, _vp, _vp, _vp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def have_ref(fma=False):
    return os.path.exists(REF_LIB_FMA if fma else REF_LIB)


def ref(fma=False):
    path = REF_LIB_FMA if fma else REF_LIB
    if path not in _refs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle ref, needs /root/reference)")
        L = C.CDLL(path)
        sig = {
            "ref_build": (_vp, [_vp, _u64, C.c_int]),
            "ref_free": (None, [_vp]),
            "ref_stats": (None, [_vp, _vp]),
            "ref_intersect": (None, [_vp, _vp, _u64, _vp, _vp]),
            "ref_intersect_avx": (C.c_double, [_vp, _vp, _u64, C.c_int, _vp, _vp]),
            "ref_export": (None, [_vp, _vp, _vp, _vp, _vp]),
            "ref_init_tris": (None, [_vp, _u64, _vp, _vp, _vp]),
            "ref_prim_rays": (None, [_vp, _vp, _u64, _u64, _u64, _vp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _refs[path] = L
    return _refs[path]


# ------------------------------------------------------------ restatement --

def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    out = np.empty(4, np.uint32)
    lib().orc_philox(_p(c), int(key[0]), int(key[1]), _p(out))
    return out


def sincos2pi(u):
    u = np.ascontiguousarray(u, np.float32)
    s, c = np.empty_like(u), np.empty_like(u)
    lib().orc_sincos2pi(_p(u), u.size, _p(s), _p(c))
    return s, c


def sample_hemisphere(normals, u1, u2):
    n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    u1 = np.ascontiguousarray(u1, np.float32)
    u2 = np.ascontiguousarray(u2, np.float32)
    out = np.empty_like(n)
    lib().orc_sample_hemisphere(_p(n), _p(u1), _p(u2), n.shape[0], _p(out))
    return out


def gen_rays(seed, box6, n):
    box6 = np.ascontiguousarray(box6, np.float32)
    out = np.empty((n, 6), np.float32)
    lib().orc_gen_rays(seed, _p(box6), n, _p(out))
    return out


def init_tris(abc):
    abc = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
    n = abc.shape[0]
    ab, ac, nn = (np.empty((n, 3), np.float32) for _ in range(3))
    lib().orc_init_tris(_p(abc), n, _p(ab), _p(ac), _p(nn))
    return ab, ac, nn


def lbvh_build(abc, leaf_size):
    """Twin of pray_build_scene: the flat layout of the GPU-built tree
    ({nodes, refs, a, ab, ac, box, max_depth}, as Tree.flat())."""
    abc = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
    n = abc.shape[0]
    nodes = np.zeros(4 * max(n, 1), np.uint32)
    refs = np.zeros(n, np.uint32)
    box = np.zeros(6, np.float32)
    depth = lib().orc_lbvh_build(_p(abc), n, leaf_size, _p(nodes), _p(refs), _p(box))
    ab, ac, _ = init_tris(abc)
    return {"nodes": nodes if n else nodes[:0], "refs": refs, "a": np.ascontiguousarray(abc[:, :3]), "ab": ab,
            "ac": ac, "box": box, "max_depth": int(depth)}


def sah_build(abc, leaf_size=8, max_depth=40, ct=2.0, ci=1.0):
    """Twin of pray_build_scene(PRAY_BUILD_SAH): the flat layout
    ({nodes, refs, a, ab, ac, box, max_depth})."""
    abc = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
    n = abc.shape[0]
    nr, depth, box = _u64(0), _u32(0), np.zeros(6, np.float32)
    L = lib()
    ne = L.orc_sah_build(_p(abc), n, leaf_size, max_depth, ct, ci, None, 0, None, 0, C.byref(nr), C.byref(depth),
                         _p(box))
    nodes = np.zeros(2 * ne, np.uint32)
    refs = np.zeros(max(nr.value, 1), np.uint32)
    L.orc_sah_build(_p(abc), n, leaf_size, max_depth, ct, ci, _p(nodes), nodes.size, _p(refs), refs.size,
                    C.byref(nr), C.byref(depth), _p(box))
    ab, ac, _ = init_tris(abc)
    return {"nodes": nodes, "refs": refs[:nr.value], "a": np.ascontiguousarray(abc[:, :3]), "ab": ab, "ac": ac,
            "box": box, "max_depth": int(depth.value)}


def mt(ray6, a, ab, ac, eps):
    t = np.zeros(1, np.float32)
    arrs = [np.ascontiguousarray(x, np.float32) for x in (ray6, a, ab, ac)]
    h = lib().orc_mt(*[_p(x) for x in arrs], eps, _p(t))
    return bool(h), float(t[0])


def mt_avx(rays, a, ab, ac, eps):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    hit = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    arrs = [np.ascontiguousarray(x, np.float32) for x in (a, ab, ac)]
    lib().orc_mt_avx(_p(rays), n, *[_p(x) for x in arrs], eps, _p(hit), _p(t))
    return hit.astype(bool), t


def slab(rays, box6):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    s, a = np.empty(n, np.int32), np.empty(n, np.int32)
    box6 = np.ascontiguousarray(box6, np.float32)
    lib().orc_slab(_p(rays), n, _p(box6), _p(s), _p(a))
    return s.astype(bool), a.astype(bool)


def bf_closest(a, ab, ac, rays, eps):
    a, ab, ac = (np.ascontiguousarray(x, np.float32).reshape(-1, 3) for x in (a, ab, ac))
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    idx, t = np.empty(n, np.int32), np.empty(n, np.float32)
    lib().orc_bf_closest(_p(a), _p(ab), _p(ac), a.shape[0], _p(rays), n, eps, _p(idx), _p(t))
    return idx, t


def ref_closest(export, a, ab, ac, rays, eps):
    """Exhaustive reference traversal (scalar) over a preorder tree export."""
    ni, nf, tr = export
    a, ab, ac = (np.ascontiguousarray(x, np.float32).reshape(-1, 3) for x in (a, ab, ac))
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    idx, t = np.empty(n, np.int32), np.empty(n, np.float32)
    lib().orc_ref_closest(_p(ni), _p(nf), ni.shape[0], _p(tr), _p(a), _p(ab), _p(ac), _p(rays), n, eps, _p(idx),
                          _p(t))
    return idx, t


def ref_closest_avx(export, a, ab, ac, rays, eps, threads):
    """AVX packet reference traversal; returns (idx, t, seconds)."""
    ni, nf, tr = export
    a, ab, ac = (np.ascontiguousarray(x, np.float32).reshape(-1, 3) for x in (a, ab, ac))
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    idx, t = np.empty(n, np.int32), np.empty(n, np.float32)
    sec = lib().orc_ref_closest_avx(_p(ni), _p(nf), ni.shape[0], _p(tr), _p(a), _p(ab), _p(ac), _p(rays), n, eps,
                                    threads, _p(idx), _p(t))
    return idx, t, sec


def flat_trace(flat, rays, mode=0, eps=1e-4, limit2=None, counts=False, boxed=None):
    """The GPU traversal's twin over the flat layout, or with boxed node steps
    over `boxed`, a boxed layout from nostrum_pray_amd.boxed_layout (width 2:
    node_levels 3; width 4: node_levels 4)."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    idx, t = np.empty(n, np.int32), np.empty(n, np.float32)
    cnt = np.empty((n, 2), np.uint32) if counts else None
    lim = None if limit2 is None else np.ascontiguousarray(limit2, np.float32)
    if boxed is not None:
        rec = np.ascontiguousarray(boxed["rec"], np.uint32)
        brefs = np.ascontiguousarray(boxed["refs"], np.uint32)
        if boxed.get("width", 2) == 4:
            lib().orc_trace_boxed4(_p(rec), boxed["root"], _p(brefs), _p(flat["a"]), _p(flat["ab"]), _p(flat["ac"]),
                                   _p(flat["box"]), flat["a"].shape[0], _p(rays), n, mode, eps, _p(lim), _p(idx),
                                   _p(t), _p(cnt))
        else:
            lib().orc_trace_boxed(_p(rec), _p(brefs), _p(flat["a"]), _p(flat["ab"]), _p(flat["ac"]), _p(flat["box"]),
                                  flat["a"].shape[0], _p(rays), n, mode, eps, _p(lim), _p(idx), _p(t), _p(cnt))
    else:
        lib().orc_flat_trace(_p(flat["nodes"]), _p(flat["refs"]), _p(flat["a"]), _p(flat["ab"]), _p(flat["ac"]),
                             _p(flat["box"]), flat["a"].shape[0], _p(rays), n, mode, eps, _p(lim), _p(idx), _p(t),
                             _p(cnt))
    return (idx, t, cnt) if counts else (idx, t)


def render_rt(flat, cam16, w, h, lights, bg, eps):
    """The reference rayTrace restated (AVX packets + scalar remainder)."""
    L = lib()
    L.orc_render_rt.restype = None
    L.orc_render_rt.argtypes = [_vp] * 9 + [_u64, _u64, _vp, C.c_int, _vp, _f32, _vp]
    lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 6)
    rgb = np.empty((h, w, 3), np.uint8)
    L.orc_render_rt(_p(flat["nodes"]), _p(flat["refs"]), _p(flat["a"]), _p(flat["ab"]), _p(flat["ac"]),
                    _p(flat["box"]), _p(flat["normal"]), _p(flat["color"]), _p(np.ascontiguousarray(cam16, np.float32)),
                    w, h, _p(lights), lights.shape[0], _p(np.ascontiguousarray(bg, np.float32)), eps, _p(rgb))
    return rgb


def render_pt(flat, cam16, w, h, spp, max_depth, rr_depth, seed, bg, eps, queues=False):
    """Path tracing restated with the build's Philox streams; (hdr, rgb), and
    with queues=True also the paths traced per bounce."""
    L = lib()
    L.orc_render_pt.restype = None
    L.orc_render_pt.argtypes = [_vp] * 10 + [_u64, _u64, _u32, _u32, _u32, _u64, _vp, _f32, _vp, _vp, _vp]
    hdr = np.empty((h, w, 3), np.float32)
    rgb = np.empty((h, w, 3), np.uint8)
    q = np.zeros(max_depth + 1, np.uint64) if queues else None
    L.orc_render_pt(_p(flat["nodes"]), _p(flat["refs"]), _p(flat["a"]), _p(flat["ab"]), _p(flat["ac"]),
                    _p(flat["box"]), _p(flat["normal"]), _p(flat["color"]), _p(flat["emit"]),
                    _p(np.ascontiguousarray(cam16, np.float32)), w, h, spp, max_depth, rr_depth, seed,
                    _p(np.ascontiguousarray(bg, np.float32)), eps, _p(hdr), _p(rgb), _p(q))
    return (hdr, rgb, q) if queues else (hdr, rgb)


def read_bmp(path):
    """24-bit BMP (as stb_image_write writes it) -> (h, w, 3) RGB, top row first."""
    data = open(path, "rb").read()
    off = int.from_bytes(data[10:14], "little")
    w = int.from_bytes(data[18:22], "little", signed=True)
    h = int.from_bytes(data[22:26], "little", signed=True)
    row = (w * 3 + 3) & ~3
    px = np.frombuffer(data, np.uint8, count=row * abs(h), offset=off).reshape(abs(h), row)[:, : w * 3]
    img = px.reshape(abs(h), w, 3)[:, :, ::-1]
    return img[::-1].copy() if h > 0 else img.copy()


def camera(pos, look, fov, w, h):
    out = np.empty(16, np.float32)
    lib().orc_camera(_p(np.ascontiguousarray(pos, np.float32)), _p(np.ascontiguousarray(look, np.float32)), fov, w,
                     h, _p(out))
    return out


def camera_rays(cam16, w, h):
    cam16 = np.ascontiguousarray(cam16, np.float32)
    out = np.empty((w * h, 6), np.float32)
    lib().orc_camera_rays(_p(cam16), w, h, _p(out))
    return out


# -------------------------------------------------------------- reference --

class RefDB:
    """The reference's KDTree / KDTreeSAH / BFSearch compiled from source."""

    def __init__(self, abc, kind, fma=False):
        self._L = ref(fma)
        abc = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
        self._h = self._L.ref_build(_p(abc), abc.shape[0], kind)

    def close(self):
        if self._h:
            self._L.ref_free(self._h)
            self._h = None

    __del__ = close

    def stats(self):
        o = np.zeros(5, np.uint64)
        self._L.ref_stats(self._h, _p(o))
        return dict(zip(["nnodes", "maxdepth", "leafNodes", "emptyLeafNodes", "triangleCount"], (int(x) for x in o)))

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        idx, t = np.empty(n, np.int32), np.empty(n, np.float32)
        self._L.ref_intersect(self._h, _p(rays), n, _p(idx), _p(t))
        return idx, t

    def intersect_avx(self, rays, threads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        idx, t = np.empty(n, np.int32), np.empty(n, np.float32)
        sec = self._L.ref_intersect_avx(self._h, _p(rays), n, threads, _p(idx), _p(t))
        return idx, t, sec

    def export(self):
        sz = np.zeros(2, np.uint64)
        self._L.ref_export(self._h, _p(sz), None, None, None)
        ni = np.empty((int(sz[0]), 4), np.int32)
        nf = np.empty((int(sz[0]), 7), np.float32)
        tr = np.empty(int(sz[1]), np.uint32)
        self._L.ref_export(self._h, None, _p(ni), _p(nf), _p(tr))
        return ni, nf, tr


def ref_init_tris(abc, fma=False):
    abc = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
    n = abc.shape[0]
    ab, ac, nn = (np.empty((n, 3), np.float32) for _ in range(3))
    ref(fma).ref_init_tris(_p(abc), n, _p(ab), _p(ac), _p(nn))
    return ab, ac, nn


def ref_prim_rays(pos, look, fov, w, h, fma=False):
    out = np.empty((w * h, 6), np.float32)
    ref(fma).ref_prim_rays(_p(np.ascontiguousarray(pos, np.float32)), _p(np.ascontiguousarray(look, np.float32)),
                           fov, w, h, _p(out))
    return out


def ref_load_scene(path, fma=False):
    """The reference's readSceneDescription (io.cpp:27-137): dict with info,
    triangles (n x 9), colours, emit flags, lights and background."""
    L = ref(fma)
    L.ref_load_scene.restype, L.ref_load_scene.argtypes = _vp, [C.c_char_p]
    L.ref_scene_free.restype, L.ref_scene_free.argtypes = None, [_vp]
    L.ref_scene_get.restype, L.ref_scene_get.argtypes = None, [_vp] * 7
    h = L.ref_load_scene(os.fsencode(path))
    try:
        info = np.zeros(16, np.float64)
        L.ref_scene_get(h, _p(info), None, None, None, None, None)
        n, nl = int(info[7]), int(info[6])
        tris = np.empty((n, 9), np.float32)
        col = np.empty((n, 3), np.float32)
        emit = np.empty(n, np.uint8)
        lights = np.empty((nl, 6), np.float32)
        bg = np.empty(3, np.float32)
        L.ref_scene_get(h, None, _p(tris), _p(col), _p(emit), _p(lights), _p(bg))
    finally:
        L.ref_scene_free(h)
    return {"mode": int(info[0]), "width": int(info[1]), "height": int(info[2]), "fov": int(info[3]),
            "max_depth": int(info[4]), "spp": int(info[5]), "tris": tris, "color": col, "emit": emit,
            "lights": lights, "background": bg, "camera_position": info[8:11].astype(np.float32),
            "camera_look": info[11:14].astype(np.float32)}


# ------------------------------------------------------ radiance cache MLP --

def mlp_encode(feats):
    feats = np.ascontiguousarray(feats, np.float32).reshape(-1, 10)
    out = np.empty((feats.shape[0], 192), np.float32)
    lib().orc_mlp_encode(_p(feats), feats.shape[0], _p(out))
    return out


def mlp_forward(params, feats, emulate=True):
    params = np.ascontiguousarray(params, np.float32)
    feats = np.ascontiguousarray(feats, np.float32).reshape(-1, 10)
    out = np.empty((feats.shape[0], 3), np.float64)
    lib().orc_mlp_forward(_p(params), _p(feats), feats.shape[0], 1 if emulate else 0, _p(out))
    return out


def mlp_grad(params, feats, targets, emulate=True, denom=None):
    """(gradient float64[30720], loss) of the relative-L2 loss; `denom`
    (n x 3) freezes the loss denominators."""
    params = np.ascontiguousarray(params, np.float32)
    feats = np.ascontiguousarray(feats, np.float32).reshape(-1, 10)
    targets = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
    den = None if denom is None else np.ascontiguousarray(denom, np.float64)
    g = np.empty(params.shape[0], np.float64)
    loss = np.zeros(1, np.float64)
    lib().orc_mlp_grad(_p(params), _p(feats), _p(targets), feats.shape[0], 1 if emulate else 0, _p(den), _p(g),
                       _p(loss))
    return g, float(loss[0])


def adam(params, m, v, grad, t, lr):
    """One Adam step in place (float32 state, as the device keeps it)."""
    lib().orc_adam(_p(params), _p(m), _p(v), _p(np.ascontiguousarray(grad, np.float64)), params.shape[0], t, lr)
