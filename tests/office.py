"""The office scene shared by the physical-mode tests: interior walls, so kept paths reflect and pass
through walls; a 9 x 9 receiver grid at mid height; two materials; a half-wave dipole at 1.8 GHz."""
import os

import numpy as np

from oracle import oracle as O

STL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stl", "office.stl")
MATS = np.array([[1.8e9, -1, -1, 2.5, 0.3, 0.03], [1.8e9, -1, -1, 6.0, 0.05, 0.005]])
KW = dict(tx=(0.835938, 4.53906, 2.5), tess=3, maxR=2, maxT=1, frequency=1.8e9, antennaType=2)


def scene_and_rx():
    """The oracle's scene and the receivers."""
    sc = O.Scene.from_stl(STL)
    v = sc.triangles()["verts"].reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    g1, g2 = np.linspace(lo[0], hi[0], 9), np.linspace(lo[1], hi[1], 9)
    return sc, np.array([[a, b, (lo[2] + hi[2]) / 2 + 0.1] for a in g1 for b in g2])
