"""Radiance cache on the GPU: the BF16 matrix-core forward pass, gradients and
Adam step against the CPU model (oracle), training on the device, and the
cache path tracer (fallback equal to the plain tracer bitwise; the trained
cache lowers the error at equal samples)."""
import numpy as np
import pytest

import nostrum_pray_amd as P
import oracle as O
from conftest import say
from test_oracle_mlp import USED, random_batch, random_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cache(ctx):
    c = P.Cache(ctx, seed=5, learning_rate=1e-2, max_batch=8192)
    yield c
    c.close()


def test_initial_params_layout(cache):
    w = cache.params()
    assert w.shape == (30720,)
    assert np.all(w[:12288].reshape(64, 192)[:, ~USED] == 0) and USED.sum() == 172
    assert np.all(w[28672:].reshape(32, 64)[3:] == 0)
    assert np.abs(w[:12288]).max() <= np.sqrt(6 / 236) and np.abs(w[:12288]).max() > 0.1


def test_forward_matches_oracle(cache):
    w = random_params(21)
    cache.set_params(w)
    f, _ = random_batch(22, 5000)
    say("cache forward 5000")
    y = cache.infer(f)
    ref = O.mlp_forward(w, f, emulate=True)
    scale = np.sqrt(np.mean(ref ** 2))
    d = np.abs(y - ref)
    say(f"forward: max {d.max() / scale:.2e} rms {np.sqrt(np.mean(d ** 2)) / scale:.2e} "
        f"over 1e-4: {np.mean(d > 1e-4 * scale):.4f} (relative to output rms {scale:.3f})")
    # FP32 sums in another order than the double-precision model: most
    # outputs agree to FP32 rounding; a rounded bf16 activation that falls on
    # the other side of a rounding boundary moves a few outputs by up to ~1%
    assert np.mean(d > 1e-4 * scale) < 0.02
    assert d.max() < 2e-2 * scale
    assert np.sqrt(np.mean(d ** 2)) < 1e-3 * scale
    assert np.array_equal(cache.params(), w)


def test_gradient_matches_oracle(cache):
    w = random_params(23)
    cache.set_params(w)
    f, t = random_batch(24, 4000)
    say("cache gradient 4000")
    g, loss = cache.gradient(f, t)
    gr, lr = O.mlp_grad(w, f, t, emulate=True)
    assert abs(loss - lr) < 1e-3 * lr
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    say(f"gradient: relative difference {rel:.2e}, loss {loss:.6f} vs {lr:.6f}")
    assert rel < 1e-2, rel
    for lo, hi in zip([0, 12288, 16384, 20480, 24576, 28672], [12288, 16384, 20480, 24576, 28672, 30720]):
        d = np.linalg.norm(g[lo:hi] - gr[lo:hi]) / np.linalg.norm(gr[lo:hi])
        assert d < 2e-2, (lo, d)
    assert np.all(g[:12288].reshape(64, 192)[:, ~USED] == 0) and np.all(g[28672:].reshape(32, 64)[3:] == 0)


def test_adam_step_matches_oracle(cache):
    w = random_params(25)
    cache.set_params(w)
    f, t = random_batch(26, 2048)
    say("cache train step")
    loss = cache.train_step(f, t)
    w1 = cache.params()
    gr, lr = O.mlp_grad(w, f, t, emulate=True)
    assert abs(loss - lr) < 1e-3 * lr
    ref = w.copy()
    O.adam(ref, np.zeros_like(w), np.zeros_like(w), gr, 1, 1e-2)
    # the first Adam step moves each weight by lr * sign(gradient): compare
    # where the gradient is clearly away from zero
    big = np.abs(gr) > 1e-3 * np.abs(gr).max()
    assert big.sum() > 10000
    assert np.allclose(w1[big], ref[big], atol=1e-6)
    assert np.all(np.abs(w1 - w) <= 1e-2 * (1 + 1e-5))
    st = cache.stats()
    assert st["steps"] == 1 and abs(st["loss"] - lr) < 1e-3 * lr


def test_training_on_device_fits_a_smooth_field(ctx):
    """A few hundred steps on a smooth target of the features lower the loss
    by an order of magnitude and generalise to held-out samples."""
    c = P.Cache(ctx, seed=9, learning_rate=1e-2, max_batch=4096)
    try:
        rng = np.random.default_rng(3)

        def field(f):
            return np.stack([1 + 0.5 * np.sin(2 * np.pi * f[:, 0]) * f[:, 7],
                             0.5 + f[:, 1] * f[:, 8],
                             0.2 + f[:, 2] ** 2 + 0.3 * f[:, 9]], 1).astype(np.float32)

        say("cache training 300 steps")
        losses = []
        for _ in range(300):
            f = rng.random((4096, 10)).astype(np.float32)
            losses.append(c.train_step(f, field(f)))
        fv = rng.random((4096, 10)).astype(np.float32)
        y = c.infer(fv)
        err = np.sqrt(np.mean((y - field(fv)) ** 2)) / np.sqrt(np.mean(field(fv) ** 2))
        assert np.mean(losses[-20:]) < 0.1 * np.mean(losses[:5]), (losses[:5], losses[-20:])
        assert err < 0.15, err
    finally:
        c.close()


def cornell_small():
    s = P.Scene.cornell(3, 20000)
    t = s.build(1)
    cam = O.camera([0, 1, 3.8], [0, 0, -1], 40, 96, 72)
    return s, t, cam


@pytest.mark.parametrize("frames,spf", [(6, 1), (2, 3), (1, 6)])
def test_fallback_equals_plain_path_tracer(ctx, frames, spf):
    """With the cache never switched on, every frame is plain path tracing
    with the plain tracer's random numbers: bitwise the same image, whether
    the 6 samples come as 6 frames of one sample, 2 of 3 or 1 of 6 (one
    training step per frame)."""
    s, t, cam = cornell_small()
    sid = ctx.upload(t)
    c = P.Cache(ctx, seed=1, max_batch=4096)
    try:
        say(f"cache fallback vs plain, {frames} frames x {spf} samples")
        bg = s.info()["background"]
        hdr_c, _, stats = c.render_pt(sid, cam, 96, 72, frames, 8, 3, 41, bg, 1e-9, train_paths=4096,
                                      cache_bounce=2, max_loss=0.0, min_steps=1, samples_per_frame=spf)
        hdr_p, _ = ctx.render_pt(sid, cam, 96, 72, 6, 8, 3, 41, bg, 1e-9, want_rgb=False)
        assert np.array_equal(hdr_c.view(np.uint32), hdr_p.view(np.uint32))
        assert np.all(stats[:, 1] == 0) and np.all(stats[:, 0] > 0)
        assert c.stats()["steps"] == frames
        with pytest.raises(P.PrayError):
            c.render_pt(sid, cam, 96, 72, 1, 8, 3, 41, bg, 1e-9, train_paths=4096, samples_per_frame=0)
    finally:
        c.close()
        ctx.free_scene(sid)


@pytest.mark.parametrize("frames,spf", [(32, 1), (8, 4)])
def test_trained_cache_lowers_error_at_equal_samples(ctx, frames, spf):
    """Against a 2048-spp plain reference, 32 samples per pixel whose paths
    end in the cache at their second hit (32 frames of one sample, or 8 of
    four) have a lower RMSE than 32 plain samples per pixel."""
    s, t, cam = cornell_small()
    sid = ctx.upload(t)
    c = P.Cache(ctx, seed=2, max_batch=1 << 14)
    try:
        bg = s.info()["background"]
        say("cache gain: reference 2048 spp")
        ref, _ = ctx.render_pt(sid, cam, 96, 72, 2048, 8, 3, 1000, bg, 1e-9, want_rgb=False)
        say("cache gain: plain 32 spp")
        plain, _ = ctx.render_pt(sid, cam, 96, 72, 32, 8, 3, 7, bg, 1e-9, want_rgb=False)
        say("cache gain: cache 32 frames")
        # warm-up frames train the cache; they are plain (cache off) until
        # 16 steps were taken
        warm, _, _ = c.render_pt(sid, cam, 96, 72, 64, 8, 3, 99, bg, 1e-9, train_paths=1 << 12, cache_bounce=1,
                                 min_steps=16)
        hdr, _, stats = c.render_pt(sid, cam, 96, 72, frames, 8, 3, 7, bg, 1e-9, train_paths=1 << 12,
                                    cache_bounce=1, min_steps=16, samples_per_frame=spf)
        assert np.all(stats[:, 1] == 1)
        rmse = lambda a: float(np.sqrt(np.mean((a - ref) ** 2)))  # noqa: E731
        say(f"rmse plain {rmse(plain):.4f} cache {rmse(hdr):.4f}")
        # fewer rays per sample (two segments instead of ~5.5) and still a
        # lower error: the cache replaces the noisy tail of every path
        assert rmse(hdr) < 0.95 * rmse(plain)
        assert abs(hdr.mean() - ref.mean()) < 0.03 * ref.mean()
    finally:
        c.close()
        ctx.free_scene(sid)


@pytest.mark.parametrize("spf", [1, 4])
def test_cache_frames_without_graphs_equal_graph_replays(ctx, spf):
    """The cache tracer's frames enqueued kernel by kernel (pray_use_graphs
    off, as profilers see them) give bitwise the images, losses and weights of
    the captured-graph replays, with the cache on after two steps."""
    s, t, cam = cornell_small()
    sid = ctx.upload(t)
    out = []
    try:
        bg = s.info()["background"]
        for graphs in (True, False):
            ctx.use_graphs(graphs)
            c = P.Cache(ctx, seed=5, max_batch=1 << 12)
            try:
                say(f"cache frames, graphs {graphs}")
                hdr, _, stats = c.render_pt(sid, cam, 96, 72, 8, 8, 3, 11, bg, 1e-9, train_paths=1 << 12,
                                            cache_bounce=1, min_steps=2, samples_per_frame=spf)
                out.append((hdr, stats, c.params()))
            finally:
                c.close()
    finally:
        ctx.use_graphs(True)
        ctx.free_scene(sid)
    (h0, s0, w0), (h1, s1, w1) = out
    # what differs, if anything: the per-frame losses and cache flags, the
    # weights (and whether they are finite), the image
    detail = (s0.tolist(), s1.tolist(), bool(np.isfinite(w0).all()), bool(np.isfinite(w1).all()),
              int(np.count_nonzero(w0 != w1)), float(h0.mean()), float(h1.mean()))
    assert s0[:, 1].any(), detail
    assert np.array_equal(s0.view(np.uint32), s1.view(np.uint32)), detail
    assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32)), detail
    assert np.array_equal(h0.view(np.uint32), h1.view(np.uint32)), detail


def test_cache_frames_ignore_stale_training_rows(ctx):
    """A frame trains on its training pixels' rows of the feature buffer;
    the rows of pixels whose paths did not reach the cache keep whatever an
    earlier call left there and must not reach the weights.  Two caches of
    one seed, whose buffers first hold a NaN batch and a finite batch (a
    gradient evaluation loads the batch and leaves the weights alone), then
    render bitwise alike, with finite weights, graphs and no graphs."""
    s, t, cam = cornell_small()
    sid = ctx.upload(t)
    bg = s.info()["background"]
    n = 1 << 12
    try:
        for graphs in (True, False):
            ctx.use_graphs(graphs)
            out = []
            for fill in (np.nan, 0.25):
                c = P.Cache(ctx, seed=5, max_batch=n)
                try:
                    say(f"stale rows {fill}, graphs {graphs}")
                    c.gradient(np.full((n, P.CACHE_FEATURES), fill, np.float32), np.zeros((n, 3), np.float32))
                    hdr, _, stats = c.render_pt(sid, cam, 96, 72, 4, 8, 3, 11, bg, 1e-9, train_paths=n,
                                                cache_bounce=1, min_steps=2, samples_per_frame=2)
                    out.append((hdr, stats, c.params()))
                finally:
                    c.close()
            (h0, s0, w0), (h1, s1, w1) = out
            assert np.isfinite(w0).all() and np.isfinite(s0).all(), (graphs, s0.tolist())
            assert s0[:, 1].any()
            assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32)), (graphs, int(np.count_nonzero(w0 != w1)))
            assert np.array_equal(s0.view(np.uint32), s1.view(np.uint32)), (graphs, s0.tolist(), s1.tolist())
            assert np.array_equal(h0.view(np.uint32), h1.view(np.uint32)), graphs
    finally:
        ctx.use_graphs(True)
        ctx.free_scene(sid)
