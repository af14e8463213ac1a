"""pray_trace_device_noalloc, the trace entry point that allocates no device
memory (its counters and stack spill live in the library's static scratch),
against pray_trace_device through the C ABI: identical hit indices and
distances on two calls in a row, the second with fewer rays, for closest and
any hit.  The first call's persistent grid is larger than the static spill
area holds, so that call runs on a cut grid."""
import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O
from conftest import first_diff, say

pytestmark = pytest.mark.gpu


def test_noalloc_trace_equals_trace_device(ctx):
    s = P.Scene.terrain(512, 2024)  # 524,288 triangles
    t = s.build(2, max_depth=40, leaf_size=8, traversal_cost=2.0)
    flat = t.flat()
    depth = t.stats()["maxdepth"]
    sid = ctx.upload(t, shading=False)
    n1, n2 = 1 << 20, 3 * (1 << 15) + 17
    # 26 resident one-wave workgroups per CU at a ring of 8 entries: the full
    # grid's spill slots (lanes x depth x 16 bytes) exceed the 64 MiB scratch
    assert 26 * 256 * 64 * depth * 16 > (64 << 20)
    bufs = []
    try:
        ctx.set_tuning(sid, layout=P.LAYOUT_PACKED, stack_lds=8, leaf_ilp=1, fetch_min=24, occupancy=100, wg_waves=1)
        box = P.scene_box_expanded(t.box())
        d_rays = ctx.alloc((n1, 6), np.float32)
        ctx.gen_rays_device(2024, box, n1, d_rays.ptr)
        lim = np.random.default_rng(3).uniform(0.0, 0.3, n1).astype(np.float32)
        d_lim = ctx.to_device(lim)
        d_ri, d_rt = ctx.alloc(n1, np.int32), ctx.alloc(n1, np.float32)
        d_ni, d_nt = ctx.alloc(n1, np.int32), ctx.alloc(n1, np.float32)
        bufs = [d_rays, d_lim, d_ri, d_rt, d_ni, d_nt]
        rays = d_rays.numpy()
        for mode in (P.TRACE_CLOSEST, P.TRACE_ANY):
            lp = d_lim.ptr if mode != P.TRACE_CLOSEST else None
            say(f"no-alloc: trace_device mode={mode} n={n1}")
            ctx.trace_device(sid, d_rays.ptr, n1, d_ri.ptr, d_rt.ptr, mode, 1e-12, d_limit2=lp)
            ctx.sync()
            ri, rt = d_ri.numpy(), d_rt.numpy()
            # the existing entry point itself against the oracle on a prefix
            oi, ot = O.flat_trace(flat, rays[:4096], mode, 1e-12, limit2=lim[:4096] if lp else None)
            if mode == P.TRACE_CLOSEST:
                assert first_diff(ri[:4096], oi) is None, first_diff(ri[:4096], oi)
            else:  # any hit: the occlusion flags
                assert np.array_equal(ri[:4096] >= 0, oi >= 0)
            for n in (n1, n2):
                d_ni.upload(np.full(n1, -7, np.int32))
                d_nt.upload(np.full(n1, -7.0, np.float32))
                say(f"no-alloc: trace_device_noalloc mode={mode} n={n}")
                ctx.trace_device_noalloc(sid, d_rays.ptr, n, d_ni.ptr, d_nt.ptr, mode, 1e-12, d_limit2=lp)
                ctx.sync()
                gi, gt = d_ni.numpy(), d_nt.numpy()
                assert first_diff(gi[:n], ri[:n]) is None, (mode, n, first_diff(gi[:n], ri[:n]))
                assert np.array_equal(gt[:n].view(np.uint32), rt[:n].view(np.uint32)), (mode, n)
                # nothing written past the n rays of the call
                assert (gi[n:] == -7).all() and (gt[n:] == -7.0).all()
            assert (ri >= 0).any() and (ri < 0).any()
        # sizes beyond the maximum are refused before anything runs
        with pytest.raises(P.PrayError):
            ctx.trace_device_noalloc(sid, d_rays.ptr, 1 << 31, d_ni.ptr, d_nt.ptr, P.TRACE_CLOSEST, 1e-12)
    finally:
        for b in bufs:
            b.free()
        ctx.free_scene(sid)


def test_noalloc_boxed_layout_equals_trace_device(ctx):
    """The same over a median tree's boxed layout (boxed node steps, whose
    stack depth is the boxed layout's, deeper than the uploaded tree's):
    closest hits equal pray_trace_device's on two calls in a row, the second
    with fewer rays."""
    s = P.Scene.random(21, 60000)
    t = s.build(0)
    flat = t.flat()
    sid = ctx.upload(t, shading=False)
    bufs = []
    try:
        assert ctx.get_tuning(sid)["node_levels"] == 4
        n1, n2 = 1 << 18, 5000
        d_rays = ctx.alloc((n1, 6), np.float32)
        ctx.gen_rays_device(77, P.scene_box_expanded(t.box()), n1, d_rays.ptr)
        d_ri, d_rt = ctx.alloc(n1, np.int32), ctx.alloc(n1, np.float32)
        d_ni, d_nt = ctx.alloc(n1, np.int32), ctx.alloc(n1, np.float32)
        bufs = [d_rays, d_ri, d_rt, d_ni, d_nt]
        say(f"no-alloc boxed: trace_device n={n1}")
        ctx.trace_device(sid, d_rays.ptr, n1, d_ri.ptr, d_rt.ptr, P.TRACE_CLOSEST, 1e-9)
        ctx.sync()
        ri, rt = d_ri.numpy(), d_rt.numpy()
        bi, _ = O.bf_closest(flat["a"], flat["ab"], flat["ac"], d_rays.numpy()[:2048], 1e-9)
        assert first_diff(ri[:2048], bi) is None
        for n in (n1, n2):
            d_ni.upload(np.full(n1, -7, np.int32))
            say(f"no-alloc boxed: trace_device_noalloc n={n}")
            ctx.trace_device_noalloc(sid, d_rays.ptr, n, d_ni.ptr, d_nt.ptr, P.TRACE_CLOSEST, 1e-9)
            ctx.sync()
            gi, gt = d_ni.numpy(), d_nt.numpy()
            assert first_diff(gi[:n], ri[:n]) is None, (n, first_diff(gi[:n], ri[:n]))
            assert np.array_equal(gt[:n].view(np.uint32), rt[:n].view(np.uint32))
            assert (gi[n:] == -7).all()
        assert (ri >= 0).any() and (ri < 0).any()
    finally:
        for b in bufs:
            b.free()
        ctx.free_scene(sid)
