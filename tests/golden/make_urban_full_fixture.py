"""Writes tests/golden/urban_L9_full_<model>.npz: the oracle's result for the urban canyon at its
full size (482 triangles, icosahedron level 9 = 2,621,442 launches, R 6, T 1, 28 GHz) on every
5th receiver of the 100 x 100 grid (2,000 receivers, offset 2), for the reference's reception
sphere (model 0) and for alpha * L / sqrt(3) (model 1).

Each receiver's kept paths and field depend only on that receiver (capture, the ancestor rule,
duplicate removal and the field sum are all per receiver), so a trace over all 10,000 receivers
restricted to these 2,000 must give exactly this fixture.  Run in the build container (about
10 minutes per model on 8 cores):

    python tests/golden/make_urban_full_fixture.py [model ...]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sbrsolver_amd import scenes  # noqa: E402

STRIDE, OFFSET = 5, 2


def fixture_path(model):
    return os.path.join(HERE, f"urban_L9_full_model{model}.npz")


def main(models):
    cfg = scenes.config2(tess=9)
    sc = O.Scene.from_soup(*cfg["soup"])
    sel = np.arange(OFFSET, len(cfg["rx"]), STRIDE, dtype=np.int32)
    for model in models:
        t = time.time()
        p = O.make_params(**dict(cfg["trace"], receptionModel=model))
        o = O.trace(sc, p, cfg["rx"][sel], cfg["materials"], nthreads=os.cpu_count())
        np.savez_compressed(
            fixture_path(model), rx_index=sel, bounces=np.int64(o.bounces),
            prx=o.prx, plaunch=o.plaunch, pnseg=o.pnseg, poff=o.poff, stri=o.stri, scop=o.scop,
            stype=o.stype, ssrc=o.ssrc, stgt=o.stgt, path_fields=o.path_fields, fields=o.fields)
        print(f"model {model}: {len(sel)} receivers, {o.np} paths, {o.bounces} bounces, "
              f"{time.time() - t:.0f} s", flush=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [0, 1])
