"""Writes tests/golden/cornell_24x24_converged.npz: the 24x24 Cornell box
frame (depth 8, directClamp 100, the C1 camera) converged with 65,536
samples per pixel accumulated in FP64 by the oracle (1,024 batches of 64
spp, sampleBatch 1..1024), plus the per-pixel variance of a 64-spp batch
mean, estimated from the same batches.

Run: python tests/golden/make_converged.py   (about a minute on 8 cores)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "reina-vk_amd")]

import oracle  # noqa: E402
from reina_rt import scenes  # noqa: E402
from reina_rt.scene import make_desc  # noqa: E402

W = H = 24
SPP, BATCHES, DEPTH = 64, 1024, 8


def main():
    sc, cam = scenes.cornell()
    orc = oracle.OracleScene(sc.build_arrays(), make_desc)
    pc = scenes.cornell_push(W, H, spp=SPP, max_bounces=DEPTH)
    t0 = time.time()
    s, sq, cnt = orc.render_sum64(pc, W, H, 1, BATCHES, use_bvh=False)
    assert (cnt == BATCHES).all()
    mean = s / BATCHES
    var_batch = np.maximum(sq / BATCHES - mean * mean, 0.0) * BATCHES / (BATCHES - 1)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cornell_24x24_converged.npz")
    np.savez_compressed(out, mean=mean, var_batch64=var_batch, spp=SPP * BATCHES, depth=DEPTH,
                        generator="tests/golden/make_converged.py")
    print("wrote %s in %.1f s" % (out, time.time() - t0))


if __name__ == "__main__":
    main()
