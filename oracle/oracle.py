"""ctypes wrapper of the FP64 oracle library (liboracle_sbr.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker.  The product (sbrsolver_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_sbr.so")
_lib = None


class Params(C.Structure):
    _fields_ = [
        ("tx", C.c_double * 3), ("txCoordSys", C.c_double * 9),
        ("maxR", C.c_int), ("maxT", C.c_int), ("maxTotal", C.c_int),
        ("freq", C.c_double), ("pt", C.c_double),
        ("useFresnel", C.c_int), ("tess", C.c_int), ("antennaType", C.c_int),
        ("receptionModel", C.c_int), ("launchStride", C.c_int), ("launchOffset", C.c_int),
        ("q1MeshIndexing", C.c_int), ("countCanonical", C.c_int), ("wallWidth", C.c_double),
        ("typedDedup", C.c_int), ("checkOcclusion", C.c_int), ("longMiss", C.c_int),
        ("physicalField", C.c_int),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle_sbr.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        vp, dp, ip, fp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_float)
        L.orc_scene_from_stl.restype = vp
        L.orc_scene_from_stl.argtypes = [C.c_char_p]
        L.orc_scene_from_soup.restype = vp
        L.orc_scene_from_soup.argtypes = [fp, fp, ip, C.c_int]
        for f in ("orc_scene_free", "orc_result_free", "orc_trees_free"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = None
        for f in ("orc_scene_ntri", "orc_scene_nnodes", "orc_scene_nverts", "orc_scene_nedges",
                  "orc_result_npaths", "orc_result_nsegs", "orc_trees_n"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = C.c_int
        for f in ("orc_result_bounces", "orc_result_node_visits", "orc_result_capture_records",
                  "orc_result_canon_nodes", "orc_result_canon_tris"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = C.c_longlong
        L.orc_result_seconds.argtypes = [vp]
        L.orc_result_seconds.restype = C.c_double
        L.orc_icosahedron.argtypes = [C.c_int, dp, C.c_longlong]
        L.orc_icosahedron.restype = C.c_longlong
        L.orc_trace.argtypes = [vp, C.POINTER(Params), dp, C.c_int, dp, C.c_int, C.c_int]
        L.orc_trace.restype = vp
        L.orc_trees.argtypes = [vp, C.POINTER(Params)]
        L.orc_trees.restype = vp
        L.orc_closest_hit.argtypes = [vp, dp, dp, dp, dp]
        L.orc_closest_hit.restype = C.c_int
        L.orc_closest_hit_brute.argtypes = [vp, dp, dp, dp, dp]
        L.orc_closest_hit_brute.restype = C.c_int
        L.orc_fresnel.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.orc_fresnel.restype = None
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def make_params(tx, tess, maxR, maxT=0, frequency=1.8e9, pt=1.0, txCoordSys=None, maxTotal=0, useFresnel=True,
                antennaType=1, receptionModel=0, launchStride=1, launchOffset=0, q1MeshIndexing=1,
                countCanonical=0, wallWidth=0.0, typedDedup=0, checkOcclusion=0, longMiss=0,
                physicalField=0):
    p = Params()
    for i in range(3):
        p.tx[i] = float(tx[i])
    cs = np.eye(3) if txCoordSys is None else np.asarray(txCoordSys, dtype=np.float64).reshape(3, 3)
    for i in range(9):
        p.txCoordSys[i] = float(cs.reshape(-1)[i])
    p.maxR, p.maxT, p.maxTotal = int(maxR), int(maxT), int(maxTotal)
    p.freq, p.pt = float(frequency), float(pt)
    p.useFresnel, p.tess, p.antennaType = int(bool(useFresnel)), int(tess), int(antennaType)
    p.receptionModel, p.launchStride, p.launchOffset = int(receptionModel), int(launchStride), int(launchOffset)
    p.q1MeshIndexing = int(q1MeshIndexing)
    p.countCanonical = int(countCanonical)
    p.wallWidth = float(wallWidth)
    p.typedDedup = int(typedDedup)
    p.checkOcclusion = int(checkOcclusion)
    p.longMiss = int(longMiss)
    p.physicalField = int(physicalField)
    return p


class Scene:
    def __init__(self, handle):
        if not handle:
            raise RuntimeError("oracle: scene construction failed")
        self.h = handle

    @classmethod
    def from_stl(cls, path):
        return cls(lib().orc_scene_from_stl(path.encode()))

    @classmethod
    def from_soup(cls, coords9, normals3, materials=None):
        c = np.ascontiguousarray(coords9, dtype=np.float32).reshape(-1, 9)
        n = np.ascontiguousarray(normals3, dtype=np.float32).reshape(-1, 3)
        m = None if materials is None else np.ascontiguousarray(materials, dtype=np.int32)
        return cls(lib().orc_scene_from_soup(_p(c, C.c_float), _p(n, C.c_float),
                                             None if m is None else _p(m, C.c_int), c.shape[0]))

    @classmethod
    def from_arrays(cls, tri, bvh):
        """A scene over given triangles (dict of sbr_scene_triangles: verts, normals, coplanar,
        material, mesh_tri_id) and a given DFS-ordered BVH (dict of sbr_scene_bvh_dfs: box, a, b, leaf,
        tri_dfs_id): the reference's traversal over a tree built elsewhere."""
        L = lib()
        L.orc_scene_from_arrays.restype = C.c_void_p
        L.orc_scene_from_arrays.argtypes = [C.c_int] + [C.c_void_p] * 5 + [C.c_int] + [C.c_void_p] * 5
        arr = lambda x, t: np.ascontiguousarray(x, dtype=t)
        keep = [arr(tri["verts"], np.float64), arr(tri["normals"], np.float64), arr(tri["coplanar"], np.int32),
                arr(tri["material"], np.int32), arr(tri["mesh_tri_id"], np.int32), arr(bvh["box"], np.float64),
                arr(bvh["a"], np.int32), arr(bvh["b"], np.int32), arr(bvh["leaf"], np.int32),
                arr(bvh["tri_dfs_id"], np.int32)]
        h = L.orc_scene_from_arrays(len(keep[2]), *[k.ctypes.data for k in keep[:5]], len(keep[6]),
                                    *[k.ctypes.data for k in keep[5:]])
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_free(self.h)
            self.h = None

    @property
    def ntri(self):
        return lib().orc_scene_ntri(self.h)

    def triangles(self):
        n = self.ntri
        cop = np.zeros(n, np.int32); mat = np.zeros(n, np.int32); mesh = np.zeros(n, np.int32)
        v = np.zeros((n, 9)); nm = np.zeros((n, 3))
        lib().orc_scene_triangles.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        lib().orc_scene_triangles(self.h, cop.ctypes.data, mat.ctypes.data, v.ctypes.data, nm.ctypes.data,
                                  mesh.ctypes.data)
        return dict(coplanar=cop, material=mat, verts=v, normals=nm, mesh_tri_id=mesh)

    def geometry(self):
        L = lib()
        nv, nt = L.orc_scene_nverts(self.h), self.ntri
        c = np.zeros(3 * nv, np.float32); t = np.zeros(3 * nt, np.uint32)
        L.orc_scene_geometry.argtypes = [C.c_void_p] * 3
        L.orc_scene_geometry(self.h, c.ctypes.data, t.ctypes.data)
        return c.reshape(-1, 3), t.reshape(-1, 3)

    def edges(self):
        L = lib()
        ne = L.orc_scene_nedges(self.h)
        pr = np.zeros(2 * ne, np.int32); cnt = np.zeros(ne, np.int32)
        L.orc_scene_edges.argtypes = [C.c_void_p] * 3
        L.orc_scene_edges(self.h, pr.ctypes.data, cnt.ctypes.data)
        return pr.reshape(-1, 2), cnt

    def bvh(self):
        L = lib()
        nn = L.orc_scene_nnodes(self.h)
        box = np.zeros((nn, 6)); ch = np.zeros((nn, 2), np.int32); no = np.zeros(nn, np.int32)
        leaf = np.zeros(self.ntri, np.int32)
        L.orc_scene_bvh.argtypes = [C.c_void_p] * 5
        L.orc_scene_bvh(self.h, box.ctypes.data, ch.ctypes.data, no.ctypes.data, leaf.ctypes.data)
        return dict(box=box, child=ch, nobj=no, leaf_tri=leaf)

    def closest_hit(self, org, d, brute=False):
        o = np.ascontiguousarray(org, np.float64); dd = np.ascontiguousarray(d, np.float64)
        pt = np.zeros(3); dist = np.zeros(1)
        f = lib().orc_closest_hit_brute if brute else lib().orc_closest_hit
        tri = f(self.h, _p(o, C.c_double), _p(dd, C.c_double), _p(pt, C.c_double), _p(dist, C.c_double))
        return tri, pt, dist[0]


def icosahedron(tess):
    n = 10 * 4 ** tess + 2
    out = np.zeros((n, 3))
    got = lib().orc_icosahedron(int(tess), _p(out, C.c_double), n)
    assert got == n
    return out


DEFAULT_MATERIALS = np.array([[1.8e9, 221.13, 0.05, 1.0, 0.0, 1e6]], dtype=np.float64)


class Result:
    """Per-receiver fields and kept paths (rx ascending, launch ascending, BFS order)."""

    def __init__(self, h):
        L = lib()
        self.bounces = L.orc_result_bounces(h)
        self.node_visits = L.orc_result_node_visits(h)
        self.canon_nodes = L.orc_result_canon_nodes(h)   # counting traversal (countCanonical=1)
        self.canon_tris = L.orc_result_canon_tris(h)
        self.capture_records = L.orc_result_capture_records(h)
        self.seconds = L.orc_result_seconds(h)
        np_, ns = L.orc_result_npaths(h), L.orc_result_nsegs(h)
        self.np, self.ns = np_, ns
        self.prx = np.zeros(np_, np.int32); self.plaunch = np.zeros(np_, np.int32)
        self.pnseg = np.zeros(np_, np.int32); self.poff = np.zeros(np_, np.int32)
        self.stri = np.zeros(ns, np.int32); self.scop = np.zeros(ns, np.int32); self.stype = np.zeros(ns, np.int32)
        self.ssrc = np.zeros((ns, 3)); self.stgt = np.zeros((ns, 3))
        L.orc_result_paths.argtypes = [C.c_void_p] * 10
        L.orc_result_paths(h, self.prx.ctypes.data, self.plaunch.ctypes.data, self.pnseg.ctypes.data,
                           self.poff.ctypes.data, self.stri.ctypes.data, self.scop.ctypes.data,
                           self.stype.ctypes.data, self.ssrc.ctypes.data, self.stgt.ctypes.data)
        pf = np.zeros((np_, 3, 2))
        L.orc_result_path_fields.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_result_path_fields(h, pf.ctypes.data)
        self.path_fields = pf[..., 0] + 1j * pf[..., 1]
        self.nrx = None
        self._h = h

    def _load_fields(self, nrx):
        f = np.zeros(6 * nrx)
        lib().orc_result_fields.argtypes = [C.c_void_p, C.c_void_p]
        lib().orc_result_fields(self._h, f.ctypes.data)
        f = f.reshape(nrx, 3, 2)
        self.fields = f[..., 0] + 1j * f[..., 1]
        lib().orc_result_free(self._h)
        self._h = None


def trace(scene, params, rx, materials=None, nthreads=0):
    rxa = np.ascontiguousarray(rx, dtype=np.float64).reshape(-1, 3)
    mats = DEFAULT_MATERIALS if materials is None else np.ascontiguousarray(materials, np.float64).reshape(-1, 6)
    h = lib().orc_trace(scene.h, C.byref(params), _p(rxa, C.c_double), rxa.shape[0], _p(mats, C.c_double),
                        mats.shape[0], int(nthreads))
    r = Result(h)
    r._load_fields(rxa.shape[0])
    return r


def trees(scene, params):
    L = lib()
    h = L.orc_trees(scene.h, C.byref(params))
    n = L.orc_trees_n(h)
    out = dict(launch=np.zeros(n, np.int32), depth=np.zeros(n, np.int32), code=np.zeros(n, np.int64),
               tri=np.zeros(n, np.int32), cop=np.zeros(n, np.int32), type=np.zeros(n, np.int32),
               src=np.zeros((n, 3)), tgt=np.zeros((n, 3)), pl=np.zeros(n))
    L.orc_trees_get.argtypes = [C.c_void_p] * 10
    L.orc_trees_get(h, *[out[k].ctypes.data for k in ("launch", "depth", "code", "tri", "cop", "type", "src",
                                                       "tgt", "pl")])
    L.orc_trees_free(h)
    return out


def set_pattern(gain_info):
    """Tabulated transmitter pattern for antennaType 3: gain_info of shape (n_theta, n_phi, 4)."""
    g = np.ascontiguousarray(gain_info, dtype=np.float32)
    L = lib()
    L.orc_set_pattern.argtypes = [C.c_void_p, C.c_int, C.c_int]
    if L.orc_set_pattern(g.ctypes.data, g.shape[0], g.shape[1]) != 0:
        raise ValueError("bad pattern table")


def pattern_eval(theta, phi, pt=1.0):
    """(E_theta, E_phi) of the tabulated pattern at (theta, phi) radians."""
    out = np.zeros(4)
    L = lib()
    L.orc_pattern_eval.argtypes = [C.c_double, C.c_double, C.c_double, C.c_void_p]
    L.orc_pattern_eval(float(theta), float(phi), float(pt), out.ctypes.data)
    return complex(out[0], out[1]), complex(out[2], out[3])


def fresnel(theta_i, rel_perm, sigma, freq, width=0.0, inf_wall=True, slab=False):
    """Reflection / transmission coefficients: the reference's (inf_wall, or its slab branch of the
    given width), or with slab=True the ITU-R P.2040 slab of the given width (extension)."""
    out = np.zeros(8)
    lib().orc_fresnel(theta_i, rel_perm, sigma, freq, width, 2 if slab else int(bool(inf_wall)), _p(out, C.c_double))
    c = out[0::2] + 1j * out[1::2]
    return dict(rTM=c[0], tTM=c[1], rTE=c[2], tTE=c[3])
