"""Python bindings of the Nostrum Pray MI355X tracer (ctypes over the C ABI).

The product is native: ``lib/libpray_hip.so`` (gfx950 kernels behind
``include/pray_hip.h``) and ``lib/libpray_host.so`` (scene loading, KD-tree
builders and flattening behind ``include/pray_host.h``).  This module only
loads those in-tree libraries and wraps their entry points for tests and the
benchmark; it raises ``NativeLibraryMissing`` when they have not been built.
There is no Python or CPU fallback for tracing.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libpray_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libpray_host.so")
CLI = os.path.join(HERE, "bin", "Pray")

PRAY_OK = 0
TRACE_CLOSEST, TRACE_ANY, TRACE_ANY_INCLUSIVE = 0, 1, 2
LAYOUT_INDEXED, LAYOUT_PACKED = 0, 1
BUILD_MORTON, BUILD_SAH = 0, 1
TUNING_FIELDS = ("layout", "stack_lds", "leaf_ilp", "fetch_min", "occupancy", "wg_waves", "leaf_batch",
                 "node_levels")
REFERENCE_DET_EPS = 1e-4  # geometry.h:12

ERRORS = {-1: "INVALID", -2: "HIP", -3: "NOMEM", -4: "NOSCENE", -5: "UNSUPPORTED", -6: "NODEVICE"}


class NativeLibraryMissing(RuntimeError):
    pass


class PrayError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pray error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


_hip = None
_host = None

_vp = C.c_void_p
_u32, _u64, _i32, _f32 = C.c_uint32, C.c_uint64, C.c_int32, C.c_float


def _p(a):
    """Pointer to a numpy array's data (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_vp)


def libraries():
    """Loads (once) and returns the (hip, host) ctypes libraries."""
    global _hip, _host
    if _hip is not None:
        return _hip, _host
    for path in (HIP_LIB, HOST_LIB):
        if not os.path.exists(path):
            raise NativeLibraryMissing(f"{path} is not built; run __graft_entry__.build()")
    hip = C.CDLL(HIP_LIB, mode=C.RTLD_LOCAL)
    host = C.CDLL(HOST_LIB, mode=C.RTLD_LOCAL)
    sig = {
        "pray_init": (C.c_int, [C.c_int, C.POINTER(_vp)]),
        "pray_destroy": (None, [_vp]),
        "pray_last_error": (C.c_char_p, [_vp]),
        "pray_mem_ok": (C.c_int, [_vp, _u64, _u64, _u64, _u64, C.POINTER(C.c_int)]),
        "pray_upload_scene": (C.c_int, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _u32,
                                        C.POINTER(C.c_int)]),
        "pray_free_scene": (C.c_int, [_vp, C.c_int]),
        "pray_set_tuning": (C.c_int, [_vp, C.c_int, _vp]),
        "pray_get_tuning": (C.c_int, [_vp, C.c_int, _vp]),
        "pray_autotune": (C.c_int, [_vp, C.c_int, _u32, C.c_int, _vp]),
        "pray_autotune_pt": (C.c_int, [_vp, C.c_int, _vp, _u32, _u32, _u32, _u32, _u64, _vp, _f32, _vp]),
        "pray_render_pt_batch": (C.c_int, [_vp, _u32, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp, _vp, _f32, _vp, _vp,
                                           C.POINTER(_f32)]),
        "pray_last_render_rays": (C.c_int, [_vp, C.POINTER(_u64)]),
        "pray_last_render_queues": (C.c_int, [_vp, _vp, _u32, C.POINTER(_u32)]),
        "pray_trace": (C.c_int, [_vp, C.c_int, _vp, _u64, C.c_int, _f32, _vp, _vp, _vp]),
        "pray_trace_device": (C.c_int, [_vp, C.c_int, _vp, _u64, C.c_int, _f32, _vp, _vp, _vp, _vp]),
        "pray_trace_counts_device": (C.c_int, [_vp, C.c_int, _vp, _u64, C.c_int, _f32, _vp, _vp, _vp, _vp, _vp]),
        "pray_trace_device_noalloc": (C.c_int, [_vp, C.c_int, _vp, _u64, C.c_int, _f32, _vp, _vp, _vp, _vp]),
        "pray_use_graphs": (C.c_int, [_vp, C.c_int]),
        "pray_set_path_budget": (C.c_int, [_vp, _u64]),
        "pray_gen_rays_device": (C.c_int, [_vp, _u64, _vp, _u64, _vp, _vp]),
        "pray_gen_rays": (C.c_int, [_vp, _u64, _vp, _u64, _vp]),
        "pray_camera_rays": (C.c_int, [_vp, _vp, _u32, _u32, _vp]),
        "pray_render_rt": (C.c_int, [_vp, C.c_int, _vp, _u32, _u32, _vp, _u32, _vp, _f32, _vp, _vp]),
        "pray_render_pt": (C.c_int, [_vp, C.c_int, _vp, _u32, _u32, _u32, _u32, _u32, _u64, _vp, _f32, _vp, _vp]),
        "pray_device_alloc": (C.c_int, [_vp, _u64, C.POINTER(_vp)]),
        "pray_device_free": (C.c_int, [_vp, _vp]),
        "pray_host_alloc": (C.c_int, [_vp, _u64, C.POINTER(_vp)]),
        "pray_host_free": (C.c_int, [_vp]),
        "pray_copy_to_device": (C.c_int, [_vp, _vp, _vp, _u64]),
        "pray_copy_to_host": (C.c_int, [_vp, _vp, _vp, _u64]),
        "pray_sync": (C.c_int, [_vp]),
        "pray_copy_device": (C.c_int, [_vp, _vp, _vp, _u64, _vp]),
        "pray_render_pt_ex": (C.c_int, [_vp, C.c_int, _vp, _u32, _u32, _u32, _u32, _u32, _u64, _vp, _f32, _vp,
                                        _vp, C.c_int]),
        "pray_cache_create": (C.c_int, [_vp, _u64, _f32, _u32, C.POINTER(C.c_int)]),
        "pray_cache_destroy": (C.c_int, [_vp, C.c_int]),
        "pray_cache_get_params": (C.c_int, [_vp, C.c_int, _vp]),
        "pray_cache_set_params": (C.c_int, [_vp, C.c_int, _vp]),
        "pray_cache_infer": (C.c_int, [_vp, C.c_int, _vp, _u64, _vp]),
        "pray_cache_gradient": (C.c_int, [_vp, C.c_int, _vp, _vp, _u64, _vp, C.POINTER(_f32)]),
        "pray_cache_train_step": (C.c_int, [_vp, C.c_int, _vp, _vp, _u64, C.POINTER(_f32)]),
        "pray_cache_stats": (C.c_int, [_vp, C.c_int, C.POINTER(_u32), C.POINTER(_f32), C.POINTER(_f32),
                                       C.POINTER(C.c_int)]),
        "pray_render_pt_cache": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _u32, _u32, _u32, _u32, _u32, _u64, _vp, _f32,
                                           _u32, _u32, _f32, _u32, _vp, _vp, _vp]),
        "pray_render_pt_cache_ex": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u64,
                                              _vp, _f32, _u32, _u32, _f32, _u32, _vp, _vp, _vp]),
        "pray_render_pt_progressive": (C.c_int, [_vp, C.c_int, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u64, _vp,
                                                 _f32, _vp, _vp, _vp]),
        "pray_build_scene": (C.c_int, [_vp, _vp, _vp, _vp, _u32, C.c_int, _u32, _u32, _f32, _f32,
                                       C.POINTER(C.c_int), C.POINTER(_f32)]),
        "pray_scene_layout": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
        "pray_boxed_layout": (C.c_int, [_vp, _u32, _vp, _u32, _vp, _vp, _vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp,
                                        _vp, _vp]),
        "pray_kernel_timing": (C.c_int, [_vp, C.c_int]),
        "pray_last_kernel_ms": (C.c_int, [_vp, C.POINTER(_f32)]),
        "pray_timed_kernels_ms": (C.c_int, [_vp, C.POINTER(_f32), C.POINTER(_u32)]),
        "pray_event_record": (C.c_int, [_vp, C.c_int]),
        "pray_event_elapsed": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(_f32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(hip, name)
        f.restype, f.argtypes = res, args
    hsig = {
        "ph_scene_terrain": (_vp, [_u32, _u64]),
        "ph_scene_spheres": (_vp, [_u64]),
        "ph_scene_cornell": (_vp, [_u64, _u32]),
        "ph_scene_random": (_vp, [_u64, _u32]),
        "ph_scene_from_tris": (_vp, [_vp, _vp, _vp, _u64]),
        "ph_scene_load_json": (_vp, [C.c_char_p, C.c_char_p, _u64]),
        "ph_scene_write_json": (C.c_int, [_vp, C.c_char_p]),
        "ph_scene_free": (None, [_vp]),
        "ph_scene_num_tris": (_u64, [_vp]),
        "ph_scene_get_tris": (None, [_vp, _vp, _vp, _vp]),
        "ph_scene_get_derived": (None, [_vp, _vp, _vp, _vp]),
        "ph_scene_info": (None, [_vp, _vp]),
        "ph_scene_set_render": (None, [_vp, _u32, _u32, _u32, _u32, C.c_int]),
        "ph_scene_get_lights": (None, [_vp, _vp]),
        "ph_camera": (None, [_vp, _vp]),
        "ph_prim_rays": (None, [_vp, _vp]),
        "ph_build": (_vp, [_vp, C.c_int]),
        "ph_build_depth": (_vp, [_vp, C.c_int, C.c_int]),
        "ph_build_ex": (_vp, [_vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]),
        "ph_tree_save": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _u64]),
        "ph_tree_load": (_vp, [C.c_char_p, C.c_char_p, _u64]),
        "ph_tree_free": (None, [_vp]),
        "ph_tree_stats": (None, [_vp, _vp]),
        "ph_tree_export": (None, [_vp, _vp, _vp, _vp, _vp]),
        "ph_flat_get": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "ph_upload": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(C.c_int)]),
        "ph_scene_fingerprint": (_u64, [_vp]),
        "ph_write_pfm": (C.c_int, [C.c_char_p, _u32, _u32, _vp]),
        "ph_read_pfm": (C.c_int, [C.c_char_p, _vp, _vp, C.c_char_p, _u64]),
        "ph_accum_save": (C.c_int, [C.c_char_p, _u32, _u32, _u32, _u64, _vp, C.c_char_p, _u64]),
        "ph_accum_load": (C.c_int, [C.c_char_p, _vp, C.POINTER(_u64), _vp, C.c_char_p, _u64]),
    }
    for name, (res, args) in hsig.items():
        f = getattr(host, name)
        f.restype, f.argtypes = res, args
    _hip, _host = hip, host
    return hip, host


def exported_symbols():
    """Names of every function declared in include/*.h, for export checks."""
    hip, host = libraries()
    return hip, host


# ----------------------------------------------------------------- host ---

class Scene:
    """A host scene (triangles, camera, lights, render settings)."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("scene construction failed")
        self._h = handle
        _, self._lib = libraries()

    @classmethod
    def terrain(cls, grid, seed):
        return cls(libraries()[1].ph_scene_terrain(grid, seed))

    @classmethod
    def spheres(cls, seed):
        return cls(libraries()[1].ph_scene_spheres(seed))

    @classmethod
    def cornell(cls, seed, target_tris=100_000):
        return cls(libraries()[1].ph_scene_cornell(seed, target_tris))

    @classmethod
    def random(cls, seed, n_tris):
        return cls(libraries()[1].ph_scene_random(seed, n_tris))

    @classmethod
    def from_triangles(cls, abc, color=None, emit=None):
        abc = np.ascontiguousarray(abc, dtype=np.float32).reshape(-1, 9)
        n = abc.shape[0]
        color = None if color is None else np.ascontiguousarray(color, dtype=np.float32).reshape(n, 3)
        emit = None if emit is None else np.ascontiguousarray(emit, dtype=np.uint8).reshape(n)
        return cls(libraries()[1].ph_scene_from_tris(_p(abc), _p(color), _p(emit), n))

    @classmethod
    def load_json(cls, path):
        err = C.create_string_buffer(1024)
        h = libraries()[1].ph_scene_load_json(os.fsencode(path), err, 1024)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    def write_json(self, path):
        if self._lib.ph_scene_write_json(self._h, os.fsencode(path)) != 0:
            raise OSError(f"cannot write {path}")

    def close(self):
        if self._h:
            self._lib.ph_scene_free(self._h)
            self._h = None

    __del__ = close

    @property
    def num_tris(self):
        return int(self._lib.ph_scene_num_tris(self._h))

    def triangles(self):
        n = self.num_tris
        abc = np.empty((n, 9), np.float32)
        col = np.empty((n, 3), np.float32)
        emit = np.empty(n, np.uint8)
        self._lib.ph_scene_get_tris(self._h, _p(abc), _p(col), _p(emit))
        return abc, col, emit

    def derived(self):
        n = self.num_tris
        ab, ac, nn = (np.empty((n, 3), np.float32) for _ in range(3))
        self._lib.ph_scene_get_derived(self._h, _p(ab), _p(ac), _p(nn))
        return ab, ac, nn

    def info(self):
        o = np.zeros(16, np.float64)
        self._lib.ph_scene_info(self._h, _p(o))
        return {
            "mode": int(o[0]), "width": int(o[1]), "height": int(o[2]), "fov": int(o[3]),
            "max_depth": int(o[4]), "spp": int(o[5]), "n_lights": int(o[6]),
            "background": o[7:10].astype(np.float32), "det_eps": np.float32(o[10]),
            "rr_depth": int(o[11]), "seed": int(o[12]), "tree": int(o[13]),
        }

    def set_render(self, width=0, height=0, spp=0, max_depth=0, rr_depth=-1):
        self._lib.ph_scene_set_render(self._h, width, height, spp, max_depth, rr_depth)

    def lights(self):
        n = self.info()["n_lights"]
        out = np.empty((n, 6), np.float32)
        if n:
            self._lib.ph_scene_get_lights(self._h, _p(out))
        return out

    def camera(self):
        out = np.empty(16, np.float32)
        self._lib.ph_camera(self._h, _p(out))
        return out

    def fingerprint(self):
        """Hash of everything that decides a path-traced pixel except the
        sample count (ph_scene_fingerprint)."""
        return int(self._lib.ph_scene_fingerprint(self._h))

    def prim_rays(self):
        inf = self.info()
        out = np.empty((inf["width"] * inf["height"], 6), np.float32)
        self._lib.ph_prim_rays(self._h, _p(out))
        return out

    def build(self, kind=-1, max_depth=20, leaf_size=0, traversal_cost=0.0, intersection_cost=0.0):
        """KD tree over the scene: 0 median, 1 SAH, -1 the reference policy.
        max_depth 20, leaf_size 32 and the SAH costs 5 / 1.5 are the
        reference's; 0 keeps them."""
        return Tree(self._lib.ph_build_ex(self._h, kind, max_depth, leaf_size, traversal_cost, intersection_cost))


class Tree:
    """A built KD tree and its flat device layout (host memory)."""

    def save(self, path):
        """Writes the flat layout to a file (ph_tree_save)."""
        err = C.create_string_buffer(512)
        if self._lib.ph_tree_save(self._h, str(path).encode(), err, 512) != 0:
            raise ValueError(err.value.decode())

    @classmethod
    def load(cls, path):
        """A tree read from a flat-layout file (no CPU tree)."""
        err = C.create_string_buffer(512)
        h = libraries()[1].ph_tree_load(str(path).encode(), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    def __init__(self, handle):
        if not handle:
            raise ValueError("tree build failed")
        self._h = handle
        _, self._lib = libraries()

    def close(self):
        if self._h:
            self._lib.ph_tree_free(self._h)
            self._h = None

    __del__ = close

    def stats(self):
        o = np.zeros(8, np.uint64)
        self._lib.ph_tree_stats(self._h, _p(o))
        keys = ["kind", "nnodes", "maxdepth", "leafNodes", "emptyLeafNodes", "triangleCount", "flat_nodes",
                "flat_refs"]
        return {k: int(v) for k, v in zip(keys, o)}

    def export(self):
        """Preorder export: (node_i32 [n,4], node_f32 [n,7], leaf tris)."""
        sz = np.zeros(2, np.uint64)
        self._lib.ph_tree_export(self._h, _p(sz), None, None, None)
        ni = np.empty((int(sz[0]), 4), np.int32)
        nf = np.empty((int(sz[0]), 7), np.float32)
        tr = np.empty(int(sz[1]), np.uint32)
        self._lib.ph_tree_export(self._h, None, _p(ni), _p(nf), _p(tr))
        return ni, nf, tr

    def box(self):
        """Root cell (min xyz, max xyz)."""
        b = np.empty(6, np.float32)
        self._lib.ph_flat_get(self._h, *([None] * 9), _p(b))
        return b

    def flat(self):
        sz = np.zeros(4, np.uint64)
        self._lib.ph_flat_get(self._h, _p(sz), *([None] * 9))
        nn, nr, nt, md = (int(x) for x in sz)
        d = {
            "nodes": np.empty(2 * nn, np.uint32), "refs": np.empty(nr, np.uint32),
            "a": np.empty((nt, 3), np.float32), "ab": np.empty((nt, 3), np.float32),
            "ac": np.empty((nt, 3), np.float32), "normal": np.empty((nt, 3), np.float32),
            "color": np.empty((nt, 3), np.float32), "emit": np.empty(nt, np.uint8),
            "box": np.empty(6, np.float32), "max_depth": md,
        }
        self._lib.ph_flat_get(self._h, None, _p(d["nodes"]), _p(d["refs"]), _p(d["a"]), _p(d["ab"]), _p(d["ac"]),
                              _p(d["normal"]), _p(d["color"]), _p(d["emit"]), _p(d["box"]))
        return d


# ------------------------------------------------------------------ GPU ---

class Context:
    """A pray_ctx on one GPU."""

    def __init__(self, device=0):
        hip, host = libraries()
        self._lib, self._host = hip, host
        h = _vp()
        rc = hip.pray_init(device, C.byref(h))
        if rc != PRAY_OK:
            raise PrayError(rc, "pray_init failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pray_destroy(self._h)
            self._h = None

    __del__ = close

    @property
    def handle(self):
        return self._h

    def _check(self, rc):
        if rc != PRAY_OK:
            raise PrayError(rc, self._lib.pray_last_error(self._h).decode())

    def set_path_budget(self, max_paths):
        """Paths the plain path tracers keep in flight (pray_set_path_budget);
        0: one sample per pass."""
        self._check(self._lib.pray_set_path_budget(self._h, int(max_paths)))

    def use_graphs(self, enable=True):
        """Path tracers replay captured hipGraphs (default) or enqueue kernel
        by kernel (pray_use_graphs); results are identical."""
        self._check(self._lib.pray_use_graphs(self._h, 1 if enable else 0))

    def mem_ok(self, n_tris, n_refs, n_nodes, n_rays):
        """Whether the device has room for a scene of these sizes and n_rays
        rays (pray_mem_ok)."""
        ok = C.c_int(0)
        self._check(self._lib.pray_mem_ok(self._h, n_tris, n_refs, n_nodes, n_rays, C.byref(ok)))
        return bool(ok.value)

    def upload(self, tree, shading=True):
        sid = C.c_int(-1)
        self._check(self._host.ph_upload(tree._h, self._h, 1 if shading else 0, C.byref(sid)))
        return sid.value

    def upload_flat(self, f, shading=True):
        sid = C.c_int(-1)
        nt = f["a"].shape[0]
        self._check(self._lib.pray_upload_scene(
            self._h, _p(f["nodes"]), len(f["nodes"]) // 2, _p(f["refs"]), len(f["refs"]), _p(f["a"]), _p(f["ab"]),
            _p(f["ac"]), _p(f["normal"]) if shading else None, _p(f["color"]) if shading else None,
            _p(f["emit"]) if shading else None, nt, _p(f["box"]), f["max_depth"], C.byref(sid)))
        return sid.value

    def free_scene(self, sid):
        self._check(self._lib.pray_free_scene(self._h, sid))

    def build_scene(self, abc, color=None, emit=None, method=None, leaf_size=8, max_depth=0, traversal_cost=0.0,
                    intersection_cost=0.0):
        """Builds the tree of raw triangles on the GPU (pray_build_scene,
        BUILD_SAH by default); returns (scene id, device build
        milliseconds).  With color (n x 3) the scene carries shading data."""
        method = BUILD_SAH if method is None else method
        abc = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
        n = abc.shape[0]
        col = None if color is None else np.ascontiguousarray(color, np.float32).reshape(n, 3)
        em = None if emit is None else np.ascontiguousarray(emit, np.uint8).reshape(n)
        sid, ms = C.c_int(-1), _f32()
        self._check(self._lib.pray_build_scene(self._h, _p(abc), _p(col), _p(em), n, method, leaf_size, max_depth,
                                               traversal_cost, intersection_cost, C.byref(sid), C.byref(ms)))
        return sid.value, ms.value

    def scene_layout(self, sid):
        """A scene's device layout (pray_scene_layout) as {nodes, refs, box,
        max_depth, n_tris}."""
        sizes = np.zeros(4, np.uint64)
        self._check(self._lib.pray_scene_layout(self._h, sid, _p(sizes), None, None, None))
        nodes = np.empty(2 * int(sizes[0]), np.uint32)
        refs = np.empty(int(sizes[1]), np.uint32)
        box = np.empty(6, np.float32)
        self._check(self._lib.pray_scene_layout(self._h, sid, _p(sizes), _p(nodes), _p(refs), _p(box)))
        return {"nodes": nodes, "refs": refs, "box": box, "max_depth": int(sizes[3]), "n_tris": int(sizes[2])}

    # device memory and timing on the context's stream
    def alloc(self, shape, dtype):
        return DeviceBuffer(self, shape, dtype)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, arr.shape, arr.dtype)
        b.upload(arr)
        return b

    def sync(self):
        self._check(self._lib.pray_sync(self._h))

    def copy_device(self, d_dst, d_src, nbytes, stream=None):
        self._check(self._lib.pray_copy_device(self._h, d_dst, d_src, nbytes, stream))

    def record(self, slot):
        self._check(self._lib.pray_event_record(self._h, slot))

    def elapsed_ms(self, slot0, slot1):
        ms = _f32()
        self._check(self._lib.pray_event_elapsed(self._h, slot0, slot1, C.byref(ms)))
        return float(ms.value)

    def set_tuning(self, sid, layout=-1, stack_lds=-1, leaf_ilp=-1, fetch_min=-1, occupancy=-1, wg_waves=-1,
                   leaf_batch=-1, node_levels=-1):
        v = np.array([layout, stack_lds, leaf_ilp, fetch_min, occupancy, wg_waves, leaf_batch, node_levels], np.int32)
        self._check(self._lib.pray_set_tuning(self._h, sid, _p(v)))

    def get_tuning(self, sid):
        v = np.zeros(len(TUNING_FIELDS), np.int32)
        self._check(self._lib.pray_get_tuning(self._h, sid, _p(v)))
        return dict(zip(TUNING_FIELDS, (int(x) for x in v)))

    def autotune(self, sid, n_rays=0, force=False):
        """Autotune the scene's traversal (or take its size bucket's setting)."""
        out = np.zeros(len(TUNING_FIELDS), np.int32)
        self._check(self._lib.pray_autotune(self._h, sid, n_rays, 1 if force else 0, _p(out)))
        return dict(zip(TUNING_FIELDS, (int(x) for x in out)))

    def autotune_pt(self, sid, cam16, width, height, max_depth, rr_depth, seed, background, det_eps):
        """Autotune the scene's traversal on path-tracing renders of this
        image and setting (pray_autotune_pt); returns the setting."""
        cam16 = np.ascontiguousarray(cam16, np.float32)
        bg = np.ascontiguousarray(background, np.float32).reshape(3)
        out = np.zeros(len(TUNING_FIELDS), np.int32)
        self._check(self._lib.pray_autotune_pt(self._h, sid, _p(cam16), width, height, max_depth, rr_depth, seed,
                                               _p(bg), det_eps, _p(out)))
        return dict(zip(TUNING_FIELDS, (int(x) for x in out)))

    def last_render_queues(self):
        """Paths traced per bounce by the last path-tracing call."""
        nb = _u32()
        self._check(self._lib.pray_last_render_queues(self._h, None, 0, C.byref(nb)))
        out = np.zeros(nb.value, np.uint64)
        self._check(self._lib.pray_last_render_queues(self._h, _p(out), nb.value, C.byref(nb)))
        return out

    def kernel_timing(self, enable=True):
        """Time the traversal kernel of later traces (pray_kernel_timing)."""
        self._check(self._lib.pray_kernel_timing(self._h, 1 if enable else 0))

    def last_kernel_ms(self):
        ms = _f32()
        self._check(self._lib.pray_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def timed_kernels_ms(self):
        """(mean ms, count) over the last traversal kernels timed, at most 64
        (pray_timed_kernels_ms)."""
        ms, cnt = _f32(), _u32()
        self._check(self._lib.pray_timed_kernels_ms(self._h, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def last_render_rays(self):
        r = _u64()
        self._check(self._lib.pray_last_render_rays(self._h, C.byref(r)))
        return r.value

    def render_pt_batch(self, sids, cams, sizes, spp, max_depth, rr_depth, seeds, backgrounds, det_eps):
        """Path traces several scenes in one wavefront; returns (list of HDR
        images, device milliseconds)."""
        n = len(sids)
        ids = np.ascontiguousarray(sids, np.int32)
        cams = np.ascontiguousarray(np.stack([np.asarray(c, np.float32) for c in cams]), np.float32)
        ws = np.ascontiguousarray([w for w, _ in sizes], np.uint32)
        hs = np.ascontiguousarray([h for _, h in sizes], np.uint32)
        sd = np.ascontiguousarray(seeds, np.uint64)
        bgs = np.ascontiguousarray(np.stack([np.asarray(b, np.float32) for b in backgrounds]), np.float32)
        outs = [np.empty((h, w, 3), np.float32) for w, h in sizes]
        ptrs = (_vp * n)(*[o.ctypes.data for o in outs])
        ms = _f32()
        self._check(self._lib.pray_render_pt_batch(self._h, n, _p(ids), _p(cams), _p(ws), _p(hs), spp, max_depth,
                                                   rr_depth, _p(sd), _p(bgs), det_eps, ptrs, None, C.byref(ms)))
        return outs, ms.value

    def trace(self, sid, rays, mode=TRACE_CLOSEST, det_eps=REFERENCE_DET_EPS, limit2=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        idx = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        lim = None if limit2 is None else np.ascontiguousarray(limit2, np.float32).reshape(n)
        self._check(self._lib.pray_trace(self._h, sid, _p(rays), n, mode, det_eps, _p(lim), _p(idx), _p(t)))
        return idx, t

    def trace_device(self, sid, d_rays, n, d_idx, d_t, mode=TRACE_CLOSEST, det_eps=REFERENCE_DET_EPS,
                     d_limit2=None, stream=None):
        self._check(self._lib.pray_trace_device(self._h, sid, d_rays, n, mode, det_eps, d_limit2, d_idx, d_t, stream))

    def trace_device_noalloc(self, sid, d_rays, n, d_idx, d_t, mode=TRACE_CLOSEST, det_eps=REFERENCE_DET_EPS,
                             d_limit2=None, stream=None):
        """trace_device with the library's static scratch: allocates no
        device memory (pray_trace_device_noalloc)."""
        self._check(self._lib.pray_trace_device_noalloc(self._h, sid, d_rays, n, mode, det_eps, d_limit2, d_idx, d_t,
                                                        stream))

    def trace_counts_device(self, sid, d_rays, n, d_idx, d_t, d_counts, mode=TRACE_CLOSEST,
                            det_eps=REFERENCE_DET_EPS, d_limit2=None, stream=None):
        self._check(self._lib.pray_trace_counts_device(self._h, sid, d_rays, n, mode, det_eps, d_limit2, d_idx, d_t,
                                                       d_counts, stream))

    def gen_rays(self, seed, box6, n):
        box6 = np.ascontiguousarray(box6, np.float32)
        out = np.empty((n, 6), np.float32)
        self._check(self._lib.pray_gen_rays(self._h, seed, _p(box6), n, _p(out)))
        return out

    def gen_rays_device(self, seed, box6, n, d_rays, stream=None):
        box6 = np.ascontiguousarray(box6, np.float32)
        self._check(self._lib.pray_gen_rays_device(self._h, seed, _p(box6), n, d_rays, stream))

    def camera_rays(self, cam16, width, height):
        cam16 = np.ascontiguousarray(cam16, np.float32)
        out = np.empty((width * height, 6), np.float32)
        self._check(self._lib.pray_camera_rays(self._h, _p(cam16), width, height, _p(out)))
        return out

    def host_array(self, shape, dtype):
        """A numpy array in page-locked host memory (pray_host_alloc), which
        the device copies into directly; freed once the array and every view
        of it are gone."""
        dtype = np.dtype(dtype)
        count = int(np.prod(shape))
        nbytes = max(1, count * dtype.itemsize)
        p = _vp()
        self._check(self._lib.pray_host_alloc(self._h, nbytes, C.byref(p)))
        buf = (C.c_uint8 * nbytes).from_address(p.value)
        weakref.finalize(buf, self._lib.pray_host_free, p.value)
        return np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)

    def render_rt(self, sid, cam16, width, height, lights, background, det_eps=REFERENCE_DET_EPS, want_hits=False,
                  out=None):
        """One ray-traced frame (pray_render_rt): the 8-bit image, and with
        want_hits the primary rays' hit indices.  `out` = (rgb, hits) arrays
        to fill instead of new ones, as a frame loop reuses its buffers."""
        cam16 = np.ascontiguousarray(cam16, np.float32)
        lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 6)
        bg = np.ascontiguousarray(background, np.float32).reshape(3)
        if out is None:
            rgb = np.empty((height, width, 3), np.uint8)
            hits = np.empty(width * height, np.int32) if want_hits else None
        else:
            rgb, hits = out
            if not (rgb.dtype == np.uint8 and rgb.size == 3 * width * height and rgb.flags.c_contiguous):
                raise ValueError("out rgb must be a contiguous uint8 array of height x width x 3")
            if want_hits and not (hits is not None and hits.dtype == np.int32 and hits.size == width * height
                                  and hits.flags.c_contiguous):
                raise ValueError("out hits must be a contiguous int32 array of width x height")
            if not want_hits:
                hits = None
        self._check(self._lib.pray_render_rt(self._h, sid, _p(cam16), width, height, _p(lights), lights.shape[0],
                                             _p(bg), det_eps, _p(rgb), _p(hits)))
        return (rgb, hits) if want_hits else rgb

    def render_pt(self, sid, cam16, width, height, spp, max_depth, rr_depth, seed, background, det_eps,
                  want_rgb=True, use_graph=True):
        cam16 = np.ascontiguousarray(cam16, np.float32)
        bg = np.ascontiguousarray(background, np.float32).reshape(3)
        hdr = np.empty((height, width, 3), np.float32)
        rgb = np.empty((height, width, 3), np.uint8) if want_rgb else None
        self._check(self._lib.pray_render_pt_ex(self._h, sid, _p(cam16), width, height, spp, max_depth, rr_depth,
                                                seed, _p(bg), det_eps, _p(hdr), _p(rgb), 1 if use_graph else 0))
        return hdr, rgb

    def render_pt_progressive(self, sid, cam16, width, height, first_sample, spp, max_depth, rr_depth, seed,
                              background, det_eps, accum=None, want_rgb=True):
        """Samples first_sample .. first_sample+spp-1 (pray_render_pt_progressive).
        accum: the running sum after first_sample samples (None when
        first_sample = 0).  Returns (hdr, rgb, new running sum)."""
        cam16 = np.ascontiguousarray(cam16, np.float32)
        bg = np.ascontiguousarray(background, np.float32).reshape(3)
        acc = np.zeros((height, width, 3), np.float32) if accum is None else \
            np.array(accum, np.float32, copy=True).reshape(height, width, 3)
        hdr = np.empty((height, width, 3), np.float32)
        rgb = np.empty((height, width, 3), np.uint8) if want_rgb else None
        self._check(self._lib.pray_render_pt_progressive(self._h, sid, _p(cam16), width, height, first_sample, spp,
                                                         max_depth, rr_depth, seed, _p(bg), det_eps, _p(acc),
                                                         _p(hdr), _p(rgb)))
        return hdr, rgb, acc


def _host_error(err):
    return ValueError(err.value.decode(errors="replace"))


def write_pfm(path, hdr):
    """Writes an (h, w, 3) float image as a Portable Float Map (ph_write_pfm)."""
    hdr = np.ascontiguousarray(hdr, np.float32)
    h, w, _ = hdr.shape
    if libraries()[1].ph_write_pfm(str(path).encode(), w, h, _p(hdr)) != 0:
        raise ValueError(f"cannot write {path}")


def read_pfm(path):
    """Reads a Portable Float Map as an (h, w, 3) float32 array (ph_read_pfm)."""
    lib = libraries()[1]
    err = C.create_string_buffer(512)
    wh = np.zeros(2, np.uint32)
    if lib.ph_read_pfm(str(path).encode(), _p(wh), None, err, 512) != 0:
        raise _host_error(err)
    out = np.empty((int(wh[1]), int(wh[0]), 3), np.float32)
    if lib.ph_read_pfm(str(path).encode(), _p(wh), _p(out), err, 512) != 0:
        raise _host_error(err)
    return out


def save_accum(path, accum, samples, fingerprint):
    """Saves a progressive render's running sum (ph_accum_save)."""
    accum = np.ascontiguousarray(accum, np.float32)
    h, w, _ = accum.shape
    err = C.create_string_buffer(512)
    if libraries()[1].ph_accum_save(str(path).encode(), w, h, samples, fingerprint, _p(accum), err, 512) != 0:
        raise _host_error(err)


def load_accum(path):
    """Loads a running sum saved by save_accum or the CLI's --accum:
    (sum (h, w, 3), samples, fingerprint)."""
    lib = libraries()[1]
    err = C.create_string_buffer(512)
    info = np.zeros(3, np.uint32)
    fp = _u64(0)
    if lib.ph_accum_load(str(path).encode(), _p(info), C.byref(fp), None, err, 512) != 0:
        raise _host_error(err)
    out = np.empty((int(info[1]), int(info[0]), 3), np.float32)
    if lib.ph_accum_load(str(path).encode(), _p(info), C.byref(fp), _p(out), err, 512) != 0:
        raise _host_error(err)
    return out, int(info[2]), fp.value


CACHE_PARAMS = 30720
CACHE_FEATURES = 10


class Cache:
    """The radiance-cache MLP of a context (pray_cache_*)."""

    def __init__(self, ctx, seed=1, learning_rate=1e-2, max_batch=1 << 16):
        self.ctx = ctx
        cid = C.c_int(-1)
        ctx._check(ctx._lib.pray_cache_create(ctx._h, seed, learning_rate, max_batch, C.byref(cid)))
        self.id = cid.value

    def close(self):
        if self.id is not None and self.ctx._h:
            self.ctx._check(self.ctx._lib.pray_cache_destroy(self.ctx._h, self.id))
        self.id = None

    def params(self):
        out = np.empty(CACHE_PARAMS, np.float32)
        self.ctx._check(self.ctx._lib.pray_cache_get_params(self.ctx._h, self.id, _p(out)))
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.shape == (CACHE_PARAMS,)
        self.ctx._check(self.ctx._lib.pray_cache_set_params(self.ctx._h, self.id, _p(w)))

    def infer(self, feats):
        feats = np.ascontiguousarray(feats, np.float32).reshape(-1, CACHE_FEATURES)
        out = np.empty((feats.shape[0], 3), np.float32)
        self.ctx._check(self.ctx._lib.pray_cache_infer(self.ctx._h, self.id, _p(feats), feats.shape[0], _p(out)))
        return out

    def gradient(self, feats, targets):
        feats = np.ascontiguousarray(feats, np.float32).reshape(-1, CACHE_FEATURES)
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        g = np.empty(CACHE_PARAMS, np.float32)
        loss = _f32()
        self.ctx._check(self.ctx._lib.pray_cache_gradient(self.ctx._h, self.id, _p(feats), _p(targets),
                                                          feats.shape[0], _p(g), C.byref(loss)))
        return g, loss.value

    def train_step(self, feats, targets):
        feats = np.ascontiguousarray(feats, np.float32).reshape(-1, CACHE_FEATURES)
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        loss = _f32()
        self.ctx._check(self.ctx._lib.pray_cache_train_step(self.ctx._h, self.id, _p(feats), _p(targets),
                                                            feats.shape[0], C.byref(loss)))
        return loss.value

    def stats(self):
        st, ll, lr, on = _u32(), _f32(), _f32(), C.c_int()
        self.ctx._check(self.ctx._lib.pray_cache_stats(self.ctx._h, self.id, C.byref(st), C.byref(ll), C.byref(lr),
                                                       C.byref(on)))
        return {"steps": st.value, "loss": ll.value, "loss_running": lr.value, "cache_on": bool(on.value)}

    def render_pt(self, sid, cam16, width, height, frames, max_depth, rr_depth, seed, background, det_eps,
                  train_paths=1 << 16, cache_bounce=2, max_loss=1e30, min_steps=1, want_rgb=False,
                  samples_per_frame=1):
        """Path tracing with this cache (pray_render_pt_cache, or
        pray_render_pt_cache_ex with samples_per_frame > 1); returns
        (hdr, rgb or None, per-frame {loss, cache on} array)."""
        cam16 = np.ascontiguousarray(cam16, np.float32)
        bg = np.ascontiguousarray(background, np.float32).reshape(3)
        hdr = np.empty((height, width, 3), np.float32)
        rgb = np.empty((height, width, 3), np.uint8) if want_rgb else None
        stats = np.empty((frames, 2), np.float32)
        if samples_per_frame == 1:
            self.ctx._check(self.ctx._lib.pray_render_pt_cache(
                self.ctx._h, sid, self.id, _p(cam16), width, height, frames, max_depth, rr_depth, seed, _p(bg),
                det_eps, train_paths, cache_bounce, max_loss, min_steps, _p(hdr), _p(rgb), _p(stats)))
        else:
            self.ctx._check(self.ctx._lib.pray_render_pt_cache_ex(
                self.ctx._h, sid, self.id, _p(cam16), width, height, frames, samples_per_frame, max_depth, rr_depth,
                seed, _p(bg), det_eps, train_paths, cache_bounce, max_loss, min_steps, _p(hdr), _p(rgb),
                _p(stats)))
        return hdr, rgb, stats


class DeviceBuffer:
    """A device allocation owned through the C ABI (pray_device_alloc)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        ctx._check(ctx._lib.pray_device_alloc(ctx._h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, self.dtype)
        assert arr.nbytes == self.nbytes
        self.ctx._check(self.ctx._lib.pray_copy_to_device(self.ctx._h, self.ptr, _p(arr), self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx._lib.pray_copy_to_host(self.ctx._h, _p(out), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx._h:
            self.ctx._lib.pray_device_free(self.ctx._h, self.ptr)
        self.ptr = None

    __del__ = free


def boxed_layout(flat, width=2):
    """The boxed layout the upload makes for boxed node steps
    (pray_boxed_layout; host code, no device): {"rec": records x 16 (width 2)
    or x 32 (width 4) uint32, "refs": leaf runs, "root": root word,
    "depth": root-to-leaf edges (width 2) or stack entries (width 4),
    "width"}.  PrayError (UNSUPPORTED) for a tree of fewer than two levels
    or whose layout would be too deep."""
    L = libraries()[0]
    nodes = np.ascontiguousarray(flat["nodes"], np.uint32).reshape(-1)
    refs = np.ascontiguousarray(flat["refs"], np.uint32)
    a, ab, ac = (np.ascontiguousarray(flat[k], np.float32) for k in ("a", "ab", "ac"))
    n_tris = a.size // 3
    n_rec, n_brefs, depth, root = _u32(), _u32(), _u32(), _u32()
    args = (_p(nodes), nodes.size // 2, _p(refs), refs.size, _p(a), _p(ab), _p(ac), n_tris, int(flat["max_depth"]),
            width, C.byref(n_rec), C.byref(n_brefs), C.byref(depth), C.byref(root))
    rc = L.pray_boxed_layout(*args, None, None)
    if rc != PRAY_OK:
        raise PrayError(rc, "no boxed layout for this tree")
    rec = np.empty((n_rec.value, 8 * width), np.uint32)
    brefs = np.empty(n_brefs.value, np.uint32)
    rc = L.pray_boxed_layout(*args, _p(rec), _p(brefs))
    if rc != PRAY_OK:
        raise PrayError(rc, "no boxed layout for this tree")
    return {"rec": rec, "refs": brefs, "depth": depth.value, "root": root.value, "width": width}


def scene_box_expanded(box6, frac=0.1):
    """Scene AABB expanded by `frac` of each extent (half per side) as the
    (lo xyz, span xyz) pair the ray generator takes."""
    lo = np.asarray(box6[:3], np.float64)
    hi = np.asarray(box6[3:], np.float64)
    ext = hi - lo
    lo2 = lo - 0.5 * frac * ext
    span = ext * (1 + frac)
    return np.concatenate([lo2, span]).astype(np.float32)
