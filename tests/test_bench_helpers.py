"""Host logic of bench.py on the CPU: the equal-RMSE interpolation, the BMP
reader, the CLI stage-time parser and the one-GPU refusal."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_equal_rmse_time_follows_inverse_square_root():
    # an unbiased estimator: rmse = c / sqrt(ms); interpolation in log-log is exact
    pts = [(ms, 3.0 / np.sqrt(ms)) for ms in (10.0, 20.0, 40.0, 80.0)]
    for target_ms in (10.0, 15.0, 33.0, 80.0):
        assert abs(bench.equal_rmse_time(pts, 3.0 / np.sqrt(target_ms)) - target_ms) < 1e-9 * target_ms
    # outside the measured range it extrapolates along the end segments
    assert abs(bench.equal_rmse_time(pts, 3.0 / np.sqrt(160.0)) - 160.0) < 1e-6
    assert abs(bench.equal_rmse_time(pts, 3.0 / np.sqrt(5.0)) - 5.0) < 1e-6
    # order of the points does not matter
    assert bench.equal_rmse_time(pts[::-1], 3.0 / np.sqrt(33.0)) == bench.equal_rmse_time(pts, 3.0 / np.sqrt(33.0))


def write_bmp(path, rgb, top_down=False):
    """A 24-bit BMP written here, independently of the reader under test."""
    h, w, _ = rgb.shape
    stride = (3 * w + 3) // 4 * 4
    rows = np.zeros((h, stride), np.uint8)
    rows[:, :3 * w] = rgb[:, :, ::-1].reshape(h, 3 * w)
    if not top_down:
        rows = rows[::-1]
    data = rows.tobytes()
    hdr = b"BM" + (54 + len(data)).to_bytes(4, "little") + b"\0\0\0\0" + (54).to_bytes(4, "little")
    info = ((40).to_bytes(4, "little") + w.to_bytes(4, "little", signed=True) +
            (-h if top_down else h).to_bytes(4, "little", signed=True) + (1).to_bytes(2, "little") +
            (24).to_bytes(2, "little") + bytes(24))
    with open(path, "wb") as f:
        f.write(hdr + info + data)


def test_bmp_reader(tmp_path):
    rng = np.random.default_rng(3)
    for w, h in ((1, 1), (5, 3), (203, 101), (8, 4)):  # odd widths have row padding
        img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        for top_down in (False, True):
            p = tmp_path / f"i{w}x{h}{int(top_down)}.bmp"
            write_bmp(p, img, top_down)
            assert np.array_equal(bench.bmp_rgb(p), img), (w, h, top_down)


def test_cli_stage_times():
    text = ("Reading scene description...\nFinished reading scene description. (3 ms)\n"
            "Preprocessing time: 26 ms\n\nRendering scene with ray tracing on the GPU...\n"
            "Finished rendering. (9 ms)\nWriting result image...\nFinished writing. (1 ms)\nTotal time: 36 ms\n")
    assert bench.cli_stage_ms(text) == {"preprocess_ms": 26, "render_ms": 9, "total_ms": 36}
    assert bench.cli_stage_ms("no timing here") == {"preprocess_ms": None, "render_ms": None, "total_ms": None}


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_pmc_traffic_matches_the_kernel_instance():
    """The committed traffic figure is used only for the traversal kernel
    instance it was measured on (layout, ring, ILP, levels, batching), not
    for another instance the autotuner may pick."""
    class Args:
        grid = 2048
    tag = "binsah-d40-leaf8-ct2"
    with open(os.path.join(ROOT, "profiles", f"trace_closest_terrain2048_{tag}_pmc.json")) as fh:
        entries = json.load(fh)["entries"]
    assert entries and all("tuning" in e and e["hbm_bytes_per_launch"] > 0 for e in entries)
    for e in entries:
        tun = dict(e["tuning"])
        assert bench.pmc_traffic(Args, tag, tun) == e["hbm_bytes_per_launch"]
        tun["fetch_min"] = 1 if tun["fetch_min"] != 1 else 2  # refill threshold: same kernel
        assert bench.pmc_traffic(Args, tag, tun) is not None
    other = dict(entries[0]["tuning"], stack_lds=13, leaf_batch=7)
    assert bench.pmc_traffic(Args, tag, other) is None
    assert bench.pmc_traffic(Args, "no-such-tree", entries[0]["tuning"]) is None


def test_parse_tuning():
    import nostrum_pray_amd as P
    assert bench.parse_tuning(P, "1,6,2,16") == {"layout": 1, "stack_lds": 6, "leaf_ilp": 2, "fetch_min": 16}
    full = bench.parse_tuning(P, "1,6,2,16,100,1,0,2")
    assert list(full) == list(P.TUNING_FIELDS) and full["node_levels"] == 2
