"""Writes tests/golden/traversal_counts.json: the work of one bounce of the urban-canyon workload
(482 triangles, icosahedron level 9, 6 reflections + 1 transmission) as counted by the oracle's
canonical closest-first, t-clipped BVH2 traversal (BVH::CanonicalCount in oracle/sbr_oracle.cpp):
boxes entered and triangles tested per bounce, over every STRIDE-th launch.

The counts define the algorithmic bytes of a bounce that bench.py reports:
  B = 32 (ray in) + 32 (segment out) + 32 * nodes + 48 * triangles.
They depend on the scene and rays only, not on the receivers or on any GPU kernel.

Run:  python tests/golden/make_traversal_counts.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

STRIDE = 1


def counts(stride=STRIDE, tess=9, nthreads=8):
    from oracle import oracle as O
    from sbrsolver_amd import scenes
    cfg = scenes.config2(tess=tess)
    sc = O.Scene.from_soup(*cfg["soup"])
    p = O.make_params(**dict(cfg["trace"], receptionModel=1, launchStride=stride, countCanonical=1))
    r = O.trace(sc, p, cfg["rx"][:1], cfg["materials"], nthreads=nthreads)
    return dict(workload=cfg["name"], launch_stride=stride, bounces=int(r.bounces),
                canon_nodes=int(r.canon_nodes), canon_tris=int(r.canon_tris),
                nodes_per_bounce=r.canon_nodes / r.bounces, tris_per_bounce=r.canon_tris / r.bounces)


def main():
    out = counts()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "traversal_counts.json")
    with open(path, "w") as f:
        json.dump({out["workload"]: out}, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
