"""Writes tests/golden/urban_L9_full_<model>.npz: the oracle's result for the urban canyon at its
full size (482 triangles, icosahedron level 9 = 2,621,442 launches, R 6, T 1, 28 GHz) on every
5th receiver of the 100 x 100 grid (2,000 receivers, offset 2), for the reference's reception
sphere (model 0) and for alpha * L / sqrt(3) (model 1).

Each receiver's kept paths and field depend only on that receiver (capture, the ancestor rule,
duplicate removal and the field sum are all per receiver), so a trace over all 10,000 receivers
restricted to these 2,000 must give exactly this fixture.  The oracle keeps every capture of the
receivers it traces in memory, so it runs on 200 receivers at a time.  Run in the build container
(about 15 minutes per model on 8 cores):

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
CHUNK = 200


def fixture_path(model):
    return os.path.join(HERE, f"urban_L9_full_model{model}.npz")


def main(models):
    cfg = scenes.config2(tess=9)
    sc = O.Scene.from_soup(*cfg["soup"])
    sel = np.arange(OFFSET, len(cfg["rx"]), STRIDE, dtype=np.int32)
    for model in models:
        t = time.time()
        p = O.make_params(**dict(cfg["trace"], receptionModel=model))
        parts = []
        for c0 in range(0, len(sel), CHUNK):   # the oracle holds every capture of its receivers at once
            o = O.trace(sc, p, cfg["rx"][sel[c0:c0 + CHUNK]], cfg["materials"], nthreads=os.cpu_count())
            parts.append((c0, o))
            print(f"model {model}: receivers {c0}..{c0 + CHUNK}: {o.np} paths, {time.time() - t:.0f} s", flush=True)
        cat = lambda f: np.concatenate([f(o) for _, o in parts])
        pnseg = cat(lambda o: o.pnseg)
        np.savez_compressed(
            fixture_path(model), rx_index=sel, bounces=np.int64(parts[0][1].bounces),
            prx=np.concatenate([o.prx + c0 for c0, o in parts]).astype(np.int32), plaunch=cat(lambda o: o.plaunch),
            pnseg=pnseg, poff=np.concatenate([[0], np.cumsum(pnseg)[:-1]]).astype(np.int32),
            stri=cat(lambda o: o.stri), scop=cat(lambda o: o.scop), stype=cat(lambda o: o.stype),
            ssrc=cat(lambda o: o.ssrc), stgt=cat(lambda o: o.stgt), path_fields=cat(lambda o: o.path_fields),
            fields=cat(lambda o: o.fields))
        print(f"model {model}: {len(sel)} receivers, {len(pnseg)} paths, {parts[0][1].bounces} bounces, "
              f"{time.time() - t:.0f} s", flush=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [0, 1])
