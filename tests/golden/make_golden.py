"""Writes the golden fixtures under tests/golden/ from the reference compiled
from source (oracle/Makefile -> oracle/_ref/).  Run in a container that has
the reference sources; the outputs are data (pixels, hit indices, distances).

  ref_images.npz  RGB renders of the example scenes by the reference CLI
                  (exact build: no FP contraction) and by its FMA build
  ref_hits.npz    primary-ray closest hits (triangle index, t) of the example
                  scenes from the reference KDTree (AVX packet path)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import nostrum_pray_amd as P  # noqa: E402
import oracle as O  # noqa: E402

RT_SCENES = ["cube_200x150.json", "cube_203x101.json", "head_fov30.json", "head_fov30_odd.json", "plant_256.json"]


def main():
    images, hits = {}, {}
    with tempfile.TemporaryDirectory() as tmp:
        for name in RT_SCENES:
            path = os.path.join(HERE, "scenes", name)
            for cli, tag in ((O.REF_CLI, "exact"), (O.REF_CLI_FMA, "fma")):
                out = os.path.join(tmp, f"{name}.{tag}.bmp")
                subprocess.run([cli, path, out], check=True, capture_output=True)
                images[f"{name}:{tag}"] = O.read_bmp(out)
            r = O.ref_load_scene(path)
            rays = O.ref_prim_rays(r["camera_position"], r["camera_look"], r["fov"], r["width"], r["height"])
            db = O.RefDB(r["tris"], 0)
            idx, t, _ = db.intersect_avx(rays, 4)
            hits[f"{name}:idx"] = idx
            hits[f"{name}:t"] = t
            print(name, images[f"{name}:exact"].shape, int((idx >= 0).sum()), "hits")
    np.savez_compressed(os.path.join(HERE, "ref_images.npz"), **images)
    np.savez_compressed(os.path.join(HERE, "ref_hits.npz"), **hits)


if __name__ == "__main__":
    main()
