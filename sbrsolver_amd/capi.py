"""ctypes binding of libsbr_amd.so (the C-ABI in include/sbr_capi.h).

There is no CPU fallback: every call runs the HIP library, and loading fails loudly when the
library has not been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsbr_amd.so")
_lib = None


class SbrError(RuntimeError):
    pass


class TraceParams(C.Structure):
    """sbr_trace_params: RTSolver::ExecuteRayTracing arguments plus extensions."""
    _fields_ = [
        ("tx", C.c_double * 3), ("txCoordSys", C.c_double * 9),
        ("maxR", C.c_int32), ("maxT", C.c_int32), ("maxTotal", C.c_int32),
        ("frequency", C.c_double), ("pt", C.c_double),
        ("useFresnel", C.c_int32), ("tess", C.c_int32), ("antennaType", C.c_int32),
        ("receptionModel", C.c_int32), ("launchStride", C.c_int32), ("launchOffset", C.c_int32),
        ("q1MeshIndexing", C.c_int32), ("exactTraversal", C.c_int32), ("countTraversal", C.c_int32),
        ("wallWidth", C.c_double), ("typedDedup", C.c_int32), ("checkOcclusion", C.c_int32),
        ("longMiss", C.c_int32), ("physicalField", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("launches", C.c_int64), ("bounces", C.c_int64), ("capture_records", C.c_int64),
        ("corrected", C.c_int64), ("winners", C.c_int64), ("max_depth", C.c_int64),
        ("ms_raygen", C.c_double), ("ms_trace", C.c_double), ("ms_capture", C.c_double),
        ("ms_field", C.c_double), ("ms_total", C.c_double),
        ("ms_k_trace", C.c_double), ("n_k_trace", C.c_int64), ("ms_k_capture", C.c_double),
        ("node_visits", C.c_int64), ("tri_tests", C.c_int64),
        ("trace_lds", C.c_int32), ("trace_blocks_per_cu", C.c_int32), ("heavy_nodes", C.c_int64),
        ("windows", C.c_int64), ("capture_tests", C.c_int64), ("box_tests_fp64", C.c_int64),
        ("capture_nodes", C.c_int64), ("wide_redone", C.c_int64), ("table_entries", C.c_int64),
        ("table_restarts", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("lds_scene", C.c_int32), ("blocks_per_cu", C.c_int32), ("window_nodes", C.c_int64),
                ("stack_lds", C.c_int32), ("oversub", C.c_int32), ("table_entries", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Material(C.Structure):
    _fields_ = [("frequency", C.c_double), ("transmissionLoss", C.c_double), ("reflectionLoss", C.c_double),
                ("relPermittivityRe", C.c_double), ("relPermittivityIm", C.c_double),
                ("relConductivity", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_queued", C.c_int32), ("hbm_budget_bytes", C.c_int64)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("n_rx", C.c_int32), ("slot", C.c_int32),
                ("launches", C.c_int64), ("bounces", C.c_int64), ("paths", C.c_int64), ("est_bytes", C.c_int64),
                ("ms_queued", C.c_double), ("ms_run", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("refused", C.c_int64),
                ("queued", C.c_int64), ("running", C.c_int64), ("bounces", C.c_int64), ("budget_bytes", C.c_int64),
                ("peak_inflight_bytes", C.c_int64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("window_s", C.c_double),
                ("req_per_s", C.c_double), ("mray_bounces_per_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SBR_PENDING = 1

# Every symbol the header declares, with its ctypes signature.
_vp, _i32, _i64, _dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
SIGNATURES = {
    "sbr_version": ([], C.c_int),
    "sbr_create": ([C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "sbr_destroy": ([_vp], C.c_int),
    "sbr_last_error": ([_vp], C.c_char_p),
    "sbr_scene_from_stl": ([_vp, C.c_char_p], C.c_int),
    "sbr_scene_from_soup": ([_vp, _vp, _vp, _vp, _i32], C.c_int),
    "sbr_scene_upload": ([_vp, _vp], C.c_int),
    "sbr_scene_info": ([_vp, _vp, _vp, _vp], C.c_int),
    "sbr_scene_triangles": ([_vp, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "sbr_scene_bvh_dfs": ([_vp, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "sbr_materials": ([_vp, _vp, _i32], C.c_int),
    "sbr_set_tx_pattern": ([_vp, _vp, _i32, _i32], C.c_int),
    "sbr_set_tuning": ([_vp, _vp], C.c_int),
    "sbr_get_tuning": ([_vp, _vp], C.c_int),
    "sbr_autotune": ([_vp, _vp, _vp, _i32, _vp], C.c_int),
    "sbr_generate_rays": ([_vp, _i32, _vp, _i64, _vp], C.c_int),
    "sbr_scene_from_soup_device": ([_vp, _vp, _vp, _vp, _i32, _vp], C.c_int),
    "sbr_closest_hit": ([_vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp], C.c_int),
    "sbr_closest_hit_noalloc": ([_vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp], C.c_int),
    "sbr_trace": ([_vp, _vp, _vp, _i32], C.c_int),
    "sbr_trace_batch": ([_vp, _vp, _i32, _vp, _i32], C.c_int),
    "sbr_path_fields": ([_vp, _vp, _vp, _i32, _i64] + [_vp] * 7 + [_i32, _vp, _vp], C.c_int),
    "sbr_fresnel_sweep": ([_vp, _i32, C.c_double, _vp, _vp, C.c_double, _i32, _vp, _vp], C.c_int),
    "sbr_result_path_tx": ([_vp, _vp, _i64], C.c_int),
    "sbr_result_stats": ([_vp, _vp], C.c_int),
    "sbr_result_fields": ([_vp, _vp, _i32], C.c_int),
    "sbr_result_path_counts": ([_vp, _vp, _vp], C.c_int),
    "sbr_result_paths": ([_vp] + [_vp] * 9, C.c_int),
    "sbr_result_path_fields": ([_vp, _vp, _i64], C.c_int),
    "sbr_result_tree_count": ([_vp, _vp], C.c_int),
    "sbr_result_trees": ([_vp] + [_vp] * 9, C.c_int),
    "sbr_server_create": ([_vp, _vp, C.POINTER(C.c_void_p)], C.c_int),
    "sbr_server_destroy": ([_vp], C.c_int),
    "sbr_server_last_error": ([_vp], C.c_char_p),
    "sbr_submit": ([_vp, _vp, _vp, _i32, C.POINTER(C.c_int64)], C.c_int),
    "sbr_poll": ([_vp, _i64, _i32, _vp], C.c_int),
    "sbr_request_error": ([_vp, _i64], C.c_char_p),
    "sbr_request_fields": ([_vp, _i64, _vp, _i32], C.c_int),
    "sbr_request_release": ([_vp, _i64], C.c_int),
    "sbr_server_get_stats": ([_vp, _vp], C.c_int),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SbrError(f"{LIB_PATH} is not built: run python sbrsolver_amd/build.py "
                           "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _a(x):
    return x.ctypes.data if x is not None else None


def make_params(tx, tess, maxR, maxT=0, frequency=1.8e9, pt=1.0, txCoordSys=None, maxTotal=0, useFresnel=True,
                antennaType=1, receptionModel=0, launchStride=1, launchOffset=0, q1MeshIndexing=1,
                exactTraversal=0, countTraversal=0, wallWidth=0.0, typedDedup=0, checkOcclusion=0,
                longMiss=0, physicalField=0, physical=False):
    """sbr_trace_params from keyword arguments (reference defaults).  physical=True switches on every
    correction of the reference's quirks at once: mesh indexing by id, the alpha*L/sqrt(3) reception
    radius, typed duplicate removal, the occlusion test, long misses and the s/p field transport
    (wallWidth stays as given: it is a wall thickness)."""
    if physical:
        q1MeshIndexing, receptionModel, typedDedup, checkOcclusion, longMiss, physicalField = 0, 1, 1, 1, 1, 1
    p = TraceParams()
    for i in range(3):
        p.tx[i] = float(tx[i])
    cs = np.eye(3) if txCoordSys is None else np.asarray(txCoordSys, dtype=np.float64).reshape(3, 3)
    for i in range(9):
        p.txCoordSys[i] = float(cs.reshape(-1)[i])
    p.maxR, p.maxT, p.maxTotal = int(maxR), int(maxT), int(maxTotal)
    p.frequency, p.pt = float(frequency), float(pt)
    p.useFresnel, p.tess, p.antennaType = int(bool(useFresnel)), int(tess), int(antennaType)
    p.receptionModel, p.launchStride, p.launchOffset = int(receptionModel), int(launchStride), int(launchOffset)
    p.q1MeshIndexing, p.exactTraversal = int(q1MeshIndexing), int(exactTraversal)
    p.countTraversal = int(countTraversal)
    p.wallWidth = float(wallWidth)
    p.typedDedup = int(typedDedup)
    p.checkOcclusion = int(checkOcclusion)
    p.longMiss = int(longMiss)
    p.physicalField = int(physicalField)
    return p


DEFAULT_MATERIALS = np.array([[1.8e9, 221.13, 0.05, 1.0, 0.0, 1e6]], dtype=np.float64)


class Result:
    """The context's last trace.  Statistics are read at once; fields, paths and trees when first
    asked for, which must be before the next trace on the same context (SbrError otherwise)."""

    def __init__(self, ctx, nrx, ntx=1):
        self.ctx, self.nrx, self.ntx = ctx, nrx, ntx
        self._gen = ctx._gen
        st = Stats()
        ctx._chk(lib().sbr_result_stats(ctx.h, C.byref(st)))
        self.stats = st.as_dict()
        self.bounces = st.bounces
        self._fields = None

    def _live(self):
        if self.ctx._gen != self._gen:
            raise SbrError("result superseded: a later trace ran on this context; read results before the next trace")

    @property
    def fields(self):
        """Complex field per (transmitter,) receiver and component, shape (nrx, 3) or (ntx, nrx, 3)."""
        if self._fields is None:
            self._live()
            n = self.ntx * self.nrx
            f = np.empty((n, 3, 2))
            self.ctx._chk(lib().sbr_result_fields(self.ctx.h, _a(f), n))
            c = f.view(np.complex128).reshape(n, 3)   # (re, im) pairs
            self._fields = c if self.ntx == 1 else c.reshape(self.ntx, self.nrx, 3)
        return self._fields

    def path_tx(self):
        self._live()
        n = self.stats["winners"]
        out = np.zeros(n, np.int32)
        self.ctx._chk(lib().sbr_result_path_tx(self.ctx.h, _a(out), n))
        return out

    def paths(self):
        self._live()
        L, h = lib(), self.ctx.h
        npth, nseg = C.c_int64(), C.c_int64()
        self.ctx._chk(L.sbr_result_path_counts(h, C.byref(npth), C.byref(nseg)))
        n, s = npth.value, nseg.value
        out = dict(prx=np.zeros(n, np.int32), plaunch=np.zeros(n, np.int32), pnseg=np.zeros(n, np.int32),
                   poff=np.zeros(n, np.int64), stri=np.zeros(s, np.int32), scop=np.zeros(s, np.int32),
                   stype=np.zeros(s, np.int32), ssrc=np.zeros((s, 3)), stgt=np.zeros((s, 3)))
        if n:
            self.ctx._chk(L.sbr_result_paths(h, *[_a(out[k]) for k in ("prx", "plaunch", "pnseg", "poff", "stri",
                                                                         "scop", "stype", "ssrc", "stgt")]))
        return out

    def path_fields(self):
        self._live()
        n = self.stats["winners"]
        f = np.zeros((n, 3, 2))
        self.ctx._chk(lib().sbr_result_path_fields(self.ctx.h, _a(f), n))
        return f[..., 0] + 1j * f[..., 1]

    def trees(self):
        self._live()
        L, h = lib(), self.ctx.h
        nn = C.c_int64()
        self.ctx._chk(L.sbr_result_tree_count(h, C.byref(nn)))
        n = nn.value
        out = dict(launch=np.zeros(n, np.int32), depth=np.zeros(n, np.int32), code=np.zeros(n, np.int64),
                   tri=np.zeros(n, np.int32), cop=np.zeros(n, np.int32), type=np.zeros(n, np.int32),
                   src=np.zeros((n, 3)), tgt=np.zeros((n, 3)), pl=np.zeros(n))
        if n:
            self.ctx._chk(L.sbr_result_trees(h, *[_a(out[k]) for k in ("launch", "depth", "code", "tri", "cop", "type",
                                                                         "src", "tgt", "pl")]))
        return out


class Context:
    """One sbr_ctx on one HIP device (not thread-safe, as the C-ABI)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().sbr_create(int(device), C.byref(h))
        if rc != 0:
            raise SbrError(f"sbr_create failed ({rc})")
        self.h = h
        self._gen = 0   # traces run on this context (a Result reads only the latest one)

    def close(self):
        if getattr(self, "h", None):
            lib().sbr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            msg = lib().sbr_last_error(self.h)
            raise SbrError(f"sbr error {rc}: {msg.decode() if msg else ''}")

    # ---- scene
    def scene_from_stl(self, path):
        self._chk(lib().sbr_scene_from_stl(self.h, path.encode()))

    def scene_from_soup(self, coords9, normals3, materials=None):
        c = np.ascontiguousarray(coords9, dtype=np.float32).reshape(-1, 9)
        n = np.ascontiguousarray(normals3, dtype=np.float32).reshape(-1, 3)
        m = None if materials is None else np.ascontiguousarray(materials, dtype=np.int32)
        self._chk(lib().sbr_scene_from_soup(self.h, _a(c), _a(n), _a(m), c.shape[0]))

    def scene_from_soup_device(self, coords9, normals3, materials=None):
        """The scene built on the device (linear BVH); returns (ms until the description, ms total)."""
        c = np.ascontiguousarray(coords9, dtype=np.float32).reshape(-1, 9)
        n = np.ascontiguousarray(normals3, dtype=np.float32).reshape(-1, 3)
        m = None if materials is None else np.ascontiguousarray(materials, dtype=np.int32)
        ms = np.zeros(2)
        self._chk(lib().sbr_scene_from_soup_device(self.h, _a(c), _a(n), _a(m), c.shape[0], _a(ms)))
        return float(ms[0]), float(ms[1])

    def scene_info(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(lib().sbr_scene_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(n_tri=a.value, n_node=b.value, n_coplanar=c.value)

    def scene_triangles(self):
        n = self.scene_info()["n_tri"]
        cop = np.zeros(n, np.int32); mat = np.zeros(n, np.int32); mesh = np.zeros(n, np.int32)
        v = np.zeros((n, 9)); nm = np.zeros((n, 3))
        self._chk(lib().sbr_scene_triangles(self.h, _a(cop), _a(mat), _a(v), _a(nm), _a(mesh)))
        return dict(coplanar=cop, material=mat, verts=v, normals=nm, mesh_tri_id=mesh)

    def scene_bvh_dfs(self):
        info = self.scene_info()
        nn, nt = info["n_node"], info["n_tri"]
        box = np.zeros((nn, 6)); a = np.zeros(nn, np.int32); b = np.zeros(nn, np.int32)
        leaf = np.zeros(nn, np.int32); tri = np.zeros(nt, np.int32)
        self._chk(lib().sbr_scene_bvh_dfs(self.h, _a(box), _a(a), _a(b), _a(leaf), _a(tri)))
        return dict(box=box, a=a, b=b, leaf=leaf, tri_dfs_id=tri)

    def materials(self, mats=None):
        m = DEFAULT_MATERIALS if mats is None else np.ascontiguousarray(mats, dtype=np.float64).reshape(-1, 6)
        self._chk(lib().sbr_materials(self.h, _a(m), m.shape[0]))

    def set_tx_pattern(self, gain_info):
        """Tabulated transmitter pattern for antennaType 3 (sbr_set_tx_pattern): gain_info of shape
        (n_theta, n_phi, 4) = (theta gain dB, theta phase deg, phi gain dB, phi phase deg) at
        theta = i * 180 / (n_theta - 1), phi = j * 360 / n_phi degrees."""
        g = np.ascontiguousarray(gain_info, dtype=np.float32)
        if g.ndim != 3 or g.shape[2] != 4:
            raise ValueError("gain_info must have shape (n_theta, n_phi, 4)")
        self._chk(lib().sbr_set_tx_pattern(self.h, _a(g), g.shape[0], g.shape[1]))

    def set_tuning(self, lds_scene=-1, blocks_per_cu=0, window_nodes=0, stack_lds=0, oversub=0, table_entries=0):
        t = Tuning(int(lds_scene), int(blocks_per_cu), int(window_nodes), int(stack_lds), int(oversub),
                   int(table_entries))
        self._chk(lib().sbr_set_tuning(self.h, C.byref(t)))

    def autotune(self, params, rx):
        """Times the traversal on a sample of this trace for each launch configuration and keeps
        the fastest (see sbr_autotune); returns the chosen tuning."""
        r = np.ascontiguousarray(rx, dtype=np.float64).reshape(-1, 3)
        t = Tuning()
        self._gen += 1   # the trials replace the context's last result
        self._chk(lib().sbr_autotune(self.h, C.byref(params), _a(r), r.shape[0], C.byref(t)))
        return t.as_dict()

    def get_tuning(self):
        t = Tuning()
        self._chk(lib().sbr_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    # ---- kernels
    def generate_rays(self, tess):
        n = 10 * 4 ** int(tess) + 2
        out = np.zeros((n, 3))
        got = C.c_int64()
        self._chk(lib().sbr_generate_rays(self.h, int(tess), _a(out), n, C.byref(got)))
        assert got.value == n
        return out

    def closest_hit(self, org, dirs, exact=0):
        """exact: 0 wide tree (winner confirmed on the reference tree), 1 reference tree in the
        reference's order, 2 reference tree with the clipping cuts; all give the reference's result."""
        o = np.ascontiguousarray(org, np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
        if o.shape[0] == 1 and d.shape[0] > 1:
            o = np.ascontiguousarray(np.repeat(o, d.shape[0], axis=0))
        n = d.shape[0]
        tri = np.zeros(n, np.int32); pt = np.zeros((n, 3)); dist = np.zeros(n)
        self._chk(lib().sbr_closest_hit(self.h, _a(o), _a(d), n, int(exact), _a(tri), _a(pt), _a(dist)))
        return tri, pt, dist

    def trace(self, params, rx):
        r = np.ascontiguousarray(rx, dtype=np.float64).reshape(-1, 3)
        self._gen += 1
        self._chk(lib().sbr_trace(self.h, C.byref(params), _a(r), r.shape[0]))
        return Result(self, r.shape[0])

    def trace_batch(self, params_list, rx):
        r = np.ascontiguousarray(rx, dtype=np.float64).reshape(-1, 3)
        arr = (TraceParams * len(params_list))(*params_list)
        self._gen += 1
        self._chk(lib().sbr_trace_batch(self.h, arr, len(params_list), _a(r), r.shape[0]))
        return Result(self, r.shape[0], len(params_list))


class Server:
    """sbr_server: coverage queries against the resident scene of `ctx` (copied at creation), run on
    `slots` concurrent engines and admitted by HBM budget.  Thread-safe."""

    def __init__(self, ctx, slots=3, hbm_budget_bytes=0, max_queued=0):
        cfg = ServerConfig(int(slots), int(max_queued), int(hbm_budget_bytes))
        h = C.c_void_p()
        ctx._chk(lib().sbr_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sbr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc < 0:
            msg = lib().sbr_server_last_error(self.h)
            raise SbrError(f"sbr error {rc}: {msg.decode() if msg else ''}")
        return rc

    def submit(self, params, rx):
        r = np.ascontiguousarray(rx, dtype=np.float64).reshape(-1, 3)
        qid = C.c_int64()
        self._chk(lib().sbr_submit(self.h, C.byref(params), _a(r), r.shape[0], C.byref(qid)))
        return qid.value

    def poll(self, qid, wait_ms=0):
        """(status, info): status 0 done, SBR_PENDING, or the failed query's negative status."""
        info = RequestInfo()
        rc = lib().sbr_poll(self.h, int(qid), int(wait_ms), C.byref(info))
        if rc < 0 and info.state != 3:   # the call failed (unknown id), not the query
            self._chk(rc)
        return rc, info.as_dict()

    def error(self, qid):
        m = lib().sbr_request_error(self.h, int(qid))
        return m.decode() if m else ""

    def fields(self, qid, n_rx):
        f = np.zeros((n_rx, 3, 2))
        self._chk(lib().sbr_request_fields(self.h, int(qid), _a(f), int(n_rx)))
        return f[..., 0] + 1j * f[..., 1]

    def release(self, qid):
        self._chk(lib().sbr_request_release(self.h, int(qid)))

    def stats(self):
        st = ServerStats()
        self._chk(lib().sbr_server_get_stats(self.h, C.byref(st)))
        return st.as_dict()
