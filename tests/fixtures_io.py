"""Parsers of the reference's committed output files (ASCII VTK / CSV written by RTSolver::Save*)."""
import re

import numpy as np

_NUM = r"[+-]?(?:\d+\.?\d*|\.\d+)(?:e[+-]?\d+)?"


def parse_complex(s):
    s = s.strip()
    assert s.endswith("i"), s
    m = re.match(rf"^({_NUM})({_NUM})i$", s)
    return complex(float(m.group(1)), float(m.group(2)))


def read_field_csv(path):
    """ElectricField.csv (RTSolver.cpp:692-717): X,Y,Z,Ex,Ey,Ez per receiver."""
    rows = open(path).read().strip().split("\n")[1:]
    xyz = np.array([[float(v) for v in r.split(",")[:3]] for r in rows])
    E = np.array([[parse_complex(v) for v in r.split(",")[3:]] for r in rows])
    return xyz, E


def _sections(path):
    return open(path).read().split("\n")


def read_raypath_vtk(path):
    """RayPath.vtk (RTSolver.cpp:510-610): per kept path its vertices, LaunchID and ReceiverID."""
    lines = _sections(path)
    i = next(k for k, l in enumerate(lines) if l.startswith("POINTS"))
    npts = int(lines[i].split()[1])
    pts = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(npts)])
    j = next(k for k, l in enumerate(lines) if l.startswith("LINES"))
    nl = int(lines[j].split()[1])
    polys = [[int(v) for v in lines[j + 1 + k].split()][1:] for k in range(nl)]
    def field(name):
        k = next(q for q, l in enumerate(lines) if l.startswith(name + " "))
        n = int(lines[k].split()[2])
        return np.array([float(lines[k + 1 + q]) for q in range(n)])
    return dict(points=pts, lines=polys, launch=field("LaunchID").astype(int), rx=field("ReceiverID").astype(int),
                eabs=field("Eabs"), edb=field("EdB"), phase=field("Phase"))


def read_rxloc_vtk(path):
    """RxLocations.vtk (RTSolver.cpp:612-690): per receiver Eabs, EdB and the phase sum."""
    lines = _sections(path)
    def field(name):
        k = next(q for q, l in enumerate(lines) if l.startswith(name + " "))
        n = int(lines[k].split()[2])
        return np.array([float(lines[k + 1 + q]) for q in range(n)])
    return dict(eabs=field("Eabs"), edb=field("EdB"), phase=field("Phase"))


def read_icosahedron_vtk(path):
    """Icosahedron.vtk (RTSolver.cpp:719-766): origin then origin + direction per launch."""
    lines = _sections(path)
    i = next(k for k, l in enumerate(lines) if l.startswith("POINTS"))
    npts = int(lines[i].split()[1])
    pts = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(npts)])
    return pts[0], pts[1:] - pts[0]


def read_faces_vtk(path):
    lines = _sections(path)
    i = next(k for k, l in enumerate(lines) if l.startswith("POINTS"))
    npts = int(lines[i].split()[1])
    pts = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(npts)])
    j = next(k for k, l in enumerate(lines) if l.startswith("POLYGONS"))
    nt = int(lines[j].split()[1])
    tris = np.array([[int(v) for v in lines[j + 1 + k].split()][1:] for k in range(nt)])
    def field(name):
        k = next(q for q, l in enumerate(lines) if l.startswith(name + " "))
        n = int(lines[k].split()[2])
        return np.array([int(lines[k + 1 + q]) for q in range(n)])
    return dict(points=pts, tris=tris, face=field("FaceID"), cop=field("CoplanarID"), mat=field("MaterialID"))


def read_edges_vtk(path):
    txt = open(path).read()
    lines = txt.replace("LINES", "\nLINES").split("\n")
    j = next(k for k, l in enumerate(lines) if l.startswith("LINES"))
    ne = int(lines[j].split()[1])
    edges = np.array([[int(v) for v in lines[j + 1 + k].split()][1:] for k in range(ne)])
    k = next(q for q, l in enumerate(lines) if l.startswith("Connection"))
    conn = np.array([int(lines[k + 1 + q]) for q in range(ne)])
    return edges, conn
