"""Writes tests/golden/traversal_counts.json: the work of one bounce, as counted by the oracle's
canonical closest-first, t-clipped BVH2 traversal (BVH::CanonicalCount in oracle/sbr_oracle.cpp):
boxes entered and triangles tested per bounce, over every STRIDE-th launch of

* the urban-canyon workload (482 triangles, icosahedron level 9, 6 reflections + 1 transmission),
  every launch;
* the dense indoor workload (206,660 triangles, level 10, reflections and transmissions up to 8
  each and 8 in total), every 64th launch (the per-bounce figures of a sample).

The counts define the algorithmic bytes of a bounce that bench.py reports:
  B = 32 (ray in) + 32 (segment out) + 32 * nodes + 48 * triangles.
They depend on the scene and rays only, not on the receivers or on any GPU kernel.

Run:  python tests/golden/make_traversal_counts.py [urban] [indoor]   (default: both; an entry
already in the file is kept unless named)
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "traversal_counts.json")


def counts(cfg, stride, nthreads=8):
    from oracle import oracle as O
    sc = O.Scene.from_soup(*cfg["soup"])
    p = O.make_params(**dict(cfg["trace"], receptionModel=1, launchStride=stride, countCanonical=1))
    r = O.trace(sc, p, cfg["rx"][:1], cfg["materials"], nthreads=nthreads)
    return dict(workload=cfg["name"], launch_stride=stride, bounces=int(r.bounces),
                canon_nodes=int(r.canon_nodes), canon_tris=int(r.canon_tris),
                nodes_per_bounce=r.canon_nodes / r.bounces, tris_per_bounce=r.canon_tris / r.bounces)


def main(which):
    from sbrsolver_amd import scenes
    out = {}
    if os.path.exists(PATH):
        with open(PATH) as f:
            out = json.load(f)
    jobs = {"urban": (lambda: scenes.config2(tess=9), 1), "indoor": (lambda: scenes.config3(tess=10), 64)}
    for name in which:
        make, stride = jobs[name]
        e = counts(make(), stride)
        out[e["workload"]] = e
        print(json.dumps(e), flush=True)
    with open(PATH, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:] or ["urban", "indoor"])
