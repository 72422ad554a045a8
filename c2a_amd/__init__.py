"""MI355X conservative-advancement CCD engine -- Python binding of the C ABI.

The engine lives in c2a_amd/lib/libc2a_amd.so (hand-written HIP for gfx950
plus the host-side builder); this module only marshals arrays through the C
ABI declared in include/c2a_amd.h.  There is no CPU fallback: if the library
is missing, loading fails with an error telling how to build it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libc2a_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "c2a_amd.h")

(OK, ERR_BAD_ARG, ERR_OOM, ERR_HIP, ERR_NO_DEVICE, ERR_UNSUPPORTED, ERR_EMPTY_MODEL,
 ERR_CACHE) = range(8)
FLAG_ITER_CAP, FLAG_TRANSLATION_ONLY, FLAG_STACK_OVERFLOW, FLAG_BAD_INPUT = 1, 2, 4, 8
NOALLOC_SCRATCH_MIB, NOALLOC_MAX_MESHES = 128, 32
ORDER_PLAIN, ORDER_LEARNED = 0, 1
NF = 32


class C2AError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("c2a status %d: %s" % (code, msg))
        self.code = code


class Query(C.Structure):
    _fields_ = [("pose", (C.c_double * 12) * 4), ("mesh", C.c_int32 * 2),
                ("seed_tri", C.c_int32 * 2)]


class Params(C.Structure):
    _fields_ = [("tol_d", C.c_double), ("tol_t", C.c_double), ("max_iter", C.c_int32),
                ("with_contacts", C.c_int32), ("contact_scale", C.c_double),
                ("contact_offset", C.c_double), ("spec_depth", C.c_int32),
                ("waves_per_query", C.c_int32), ("replay_log", C.c_int32), ("order", C.c_int32),
                ("scorer", C.c_void_p), ("sample_rate", C.c_double), ("sample_cap", C.c_int64),
                ("miss_depth", C.c_int32), ("ext_levels", C.c_int32),
                ("translation", C.c_int32), ("strict", C.c_int32),
                ("tail_waves", C.c_int32), ("tail_budget", C.c_int64)]


class Result(C.Structure):
    _fields_ = [("toc", C.c_double), ("distance", C.c_double), ("mint", C.c_double),
                ("R_toc", C.c_double * 9), ("T_toc", C.c_double * 3), ("p1", C.c_double * 3),
                ("p2", C.c_double * 3), ("bv_tests", C.c_int64), ("tri_tests", C.c_int64),
                ("numCA", C.c_int32), ("collisionfree", C.c_int32), ("num_contact", C.c_int32),
                ("last_tri", C.c_int32 * 2), ("flags", C.c_int32), ("R", C.c_double * 9),
                ("T", C.c_double * 3), ("abs_err", C.c_double), ("rel_err", C.c_double),
                ("UpboundTOC", C.c_double)]


class StepQuery(C.Structure):
    _fields_ = [("R1", C.c_double * 9), ("T1", C.c_double * 3), ("R2", C.c_double * 9),
                ("T2", C.c_double * 3), ("pose_s", (C.c_double * 12) * 2),
                ("pose_e", (C.c_double * 12) * 2), ("mesh", C.c_int32 * 2),
                ("seed_tri", C.c_int32 * 2), ("last_tri", C.c_int32 * 2), ("numCA", C.c_int32),
                ("pad", C.c_int32), ("mint_prev", C.c_double), ("UpboundTOC", C.c_double)]


class StepResult(C.Structure):
    _fields_ = [("distance", C.c_double), ("mint", C.c_double), ("abs_err", C.c_double),
                ("rel_err", C.c_double), ("p1", C.c_double * 3), ("p2", C.c_double * 3),
                ("bv_tests", C.c_int64), ("tri_tests", C.c_int64), ("last_tri", C.c_int32 * 2),
                ("flags", C.c_int32), ("pad", C.c_int32), ("R", C.c_double * 9),
                ("T", C.c_double * 3)]


RESULT_DTYPE = np.dtype([("toc", "f8"), ("distance", "f8"), ("mint", "f8"), ("R_toc", "f8", 9),
                         ("T_toc", "f8", 3), ("p1", "f8", 3), ("p2", "f8", 3),
                         ("bv_tests", "i8"), ("tri_tests", "i8"), ("numCA", "i4"),
                         ("collisionfree", "i4"), ("num_contact", "i4"), ("last_tri", "i4", 2),
                         ("flags", "i4"), ("R", "f8", 9), ("T", "f8", 3), ("abs_err", "f8"),
                         ("rel_err", "f8"), ("UpboundTOC", "f8")])
assert RESULT_DTYPE.itemsize == C.sizeof(Result)

# c2a_contact
CONTACT_DTYPE = np.dtype([("tri", "i4", 2), ("type", "i4", 2), ("fid", "i4", (2, 3)),
                          ("pa", "f8", 3), ("pb", "f8", 3), ("distance", "f8")])

# c2a_distance_result
DISTANCE_DTYPE = np.dtype([("distance", "f8"), ("p1", "f8", 3), ("p2", "f8", 3),
                           ("bv_tests", "i8"), ("tri_tests", "i8"), ("last_tri", "i4", 2),
                           ("flags", "i4"), ("pad", "i4")])

# c2a_static_result
STATIC_DTYPE = np.dtype([("distance", "f8"), ("p1", "f8", 3), ("p2", "f8", 3),
                         ("bv_tests", "i8"), ("tri_tests", "i8"), ("tri", "i4", 2),
                         ("found", "i4"), ("flags", "i4")])

_lib = None

# Every extern "C" entry point of include/c2a_amd.h.
EXPORTS = ["c2a_params_default", "c2a_mesh_create", "c2a_mesh_destroy", "c2a_mesh_info",
           "c2a_mesh_tri_ids", "c2a_mesh_export", "c2a_ccd_batch", "c2a_batch_create",
           "c2a_batch_run", "c2a_batch_fetch", "c2a_batch_destroy", "c2a_batch_last_ms",
           "c2a_ccd_step", "c2a_contact_count", "c2a_distance", "c2a_collide_distance",
           "c2a_motion_prepare", "c2a_motion_pose",
           "c2a_status_string", "c2a_last_error", "c2a_scorer_create", "c2a_scorer_destroy",
           "c2a_scorer_set_params", "c2a_scorer_get_params", "c2a_scorer_infer",
           "c2a_scorer_grad", "c2a_scorer_train", "c2a_scorer_save", "c2a_scorer_load",
           "c2a_batch_samples", "c2a_ccd_batch_noalloc", "c2a_contact_list",
           "c2a_batch_run_graph", "c2a_distance_queue", "c2a_distance_contact_space",
           "c2a_tolerance", "c2a_collide_pairs", "c2a_mesh_create_device", "c2a_mesh_save",
           "c2a_mesh_load", "c2a_batch_tail_count", "c2a_mesh_refit"]


def lib():
    """Loads libc2a_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("c2a_amd native library missing at %s; build it with "
                           "`python -c 'import __graft_entry__ as g; g.build()'` or "
                           "`python c2a_amd/build.py`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, dp, ip, i64 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int64
    L.c2a_params_default.argtypes = [C.POINTER(Params)]
    L.c2a_params_default.restype = None
    L.c2a_mesh_create.argtypes = [dp, ip, ip, i64, C.POINTER(vp)]
    L.c2a_mesh_destroy.argtypes = [vp]
    L.c2a_mesh_create_device.argtypes = [dp, ip, ip, i64, vp, C.POINTER(vp)]
    L.c2a_mesh_save.argtypes = [vp, C.c_char_p]
    L.c2a_mesh_refit.argtypes = [vp, dp, C.c_int32, vp]
    L.c2a_mesh_load.argtypes = [C.c_char_p, dp, ip, ip, i64, C.POINTER(vp)]
    L.c2a_mesh_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), ip]
    L.c2a_mesh_tri_ids.argtypes = [vp, ip]
    L.c2a_mesh_export.argtypes = [vp, dp, dp]
    L.c2a_ccd_batch.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(Query), i64,
                                C.POINTER(Params), vp, vp]
    L.c2a_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(Query), i64,
                                   C.POINTER(Params), C.POINTER(vp)]
    L.c2a_batch_run.argtypes = [vp, vp]
    L.c2a_batch_run_graph.argtypes = [vp, vp]
    L.c2a_batch_fetch.argtypes = [vp, vp, vp]
    L.c2a_batch_destroy.argtypes = [vp]
    L.c2a_batch_last_ms.argtypes = [vp, C.POINTER(C.c_float)]
    L.c2a_batch_tail_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.c2a_batch_samples.argtypes = [vp, vp, vp, vp, i64, C.POINTER(i64), C.POINTER(i64)]
    L.c2a_scorer_save.argtypes = [vp, C.c_char_p]
    L.c2a_scorer_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.c2a_ccd_step.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(StepQuery), i64,
                               C.POINTER(StepResult), vp]
    L.c2a_contact_count.argtypes = [C.POINTER(vp), C.c_int32, dp, ip, dp, i64, ip, vp]
    L.c2a_contact_list.argtypes = [C.POINTER(vp), C.c_int32, dp, ip, dp, i64, C.c_int32, vp,
                                   ip, vp]
    L.c2a_distance.argtypes = [C.POINTER(vp), C.c_int32, dp, ip, ip, C.c_double, C.c_double,
                               i64, vp, vp]
    L.c2a_collide_distance.argtypes = L.c2a_distance.argtypes
    L.c2a_distance_queue.argtypes = [C.POINTER(vp), C.c_int32, dp, ip, ip, C.c_double,
                                     C.c_double, C.c_int32, i64, vp, vp]
    L.c2a_distance_contact_space.argtypes = [C.POINTER(vp), C.c_int32, dp, ip, ip, dp, i64, vp,
                                             vp]
    L.c2a_tolerance.argtypes = [C.POINTER(vp), C.c_int32, dp, ip, dp, C.c_int32, i64, vp, vp]
    L.c2a_collide_pairs.argtypes = [C.POINTER(vp), C.c_int32, dp, ip, C.c_int32, i64,
                                    C.c_int32, ip, vp, vp]
    L.c2a_ccd_batch_noalloc.argtypes = [C.POINTER(vp), C.c_int32, vp, vp, vp, i64,
                                        C.POINTER(Params), vp, vp]
    L.c2a_motion_prepare.argtypes = [dp, dp, dp]
    L.c2a_motion_pose.argtypes = [dp, C.c_double, dp, dp]
    L.c2a_status_string.argtypes = [C.c_int]
    L.c2a_status_string.restype = C.c_char_p
    L.c2a_last_error.argtypes = []
    L.c2a_last_error.restype = C.c_char_p
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is C.c_int or name not in ("c2a_params_default", "c2a_status_string",
                                                "c2a_last_error"):
            if name not in ("c2a_params_default", "c2a_status_string", "c2a_last_error"):
                f.restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise C2AError(rc, lib().c2a_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def default_params(**kw):
    """c2a_params_default plus overrides; scorer= takes a Scorer (learned order)."""
    p = Params()
    lib().c2a_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "scorer":
            v = None if v is None else v.handle
        setattr(p, k, v)
    return p


class Mesh:
    """A triangle soup with its RSS hierarchy (C2A_Model after EndModel).

    device=True builds the hierarchy on the GPU (the same nodes as the host
    build).  cache=path reuses a hierarchy file written for the same
    triangles, ids and vertex indices, and otherwise builds and writes one."""

    def __init__(self, tris, ids=None, vidx=None, device=False, cache=None):
        tris = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
        self.ntri = tris.shape[0]
        ida = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        vda = None if vidx is None else np.ascontiguousarray(vidx, dtype=np.int32).reshape(-1)
        args = (_dp(tris), None if ida is None else _ip(ida), None if vda is None else _ip(vda),
                self.ntri)
        h = C.c_void_p()
        self.from_cache = False
        if cache is not None and os.path.exists(cache):
            rc = lib().c2a_mesh_load(os.fsencode(cache), *args, C.byref(h))
            if rc == OK:
                self.from_cache = True
            elif rc != ERR_CACHE:
                _check(rc)
        if not self.from_cache:
            if device:
                _check(lib().c2a_mesh_create_device(*args, None, C.byref(h)))
            else:
                _check(lib().c2a_mesh_create(*args, C.byref(h)))
            if cache is not None:
                _check(lib().c2a_mesh_save(h, os.fsencode(cache)))
        self.handle = h
        n, b, d = C.c_int64(), C.c_int64(), C.c_int32()
        _check(lib().c2a_mesh_info(h, C.byref(n), C.byref(b), C.byref(d)))
        self.nbv, self.max_depth = b.value, d.value

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().c2a_mesh_destroy(self.handle)
            except Exception:
                pass
            self.handle = None

    def export(self):
        """(nodes (nbv, 32) float64, tris (ntri, 9) float64) in build order."""
        nodes = np.zeros((self.nbv, 32))
        tris = np.zeros((self.ntri, 9))
        _check(lib().c2a_mesh_export(self.handle, _dp(nodes), _dp(tris)))
        return nodes, tris

    def refit(self, tris, device=True):
        """Refits every node's volume to moved triangles (c2a_mesh_refit): the
        same triangles, in the order given at creation, n x 9."""
        tris = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
        if tris.shape[0] != self.ntri:
            raise ValueError("refit takes the mesh's %d triangles" % self.ntri)
        _check(lib().c2a_mesh_refit(self.handle, _dp(tris), 1 if device else 0, None))

    def save(self, path):
        """Writes the hierarchy file (c2a_mesh_save)."""
        _check(lib().c2a_mesh_save(self.handle, os.fsencode(path)))

    def tri_ids(self):
        ids = np.zeros(self.ntri, dtype=np.int32)
        _check(lib().c2a_mesh_tri_ids(self.handle, _ip(ids)))
        return ids


def _mesh_table(meshes):
    arr = (C.c_void_p * len(meshes))(*[m.handle for m in meshes])
    return arr


def make_queries(P00, P01, P10, P11, mesh_pairs=None, seeds=None):
    """Packs (nq,12) pose arrays into a ctypes Query array."""
    nq = len(P00)
    qs = (Query * nq)()
    buf = np.ctypeslib.as_array(qs)
    poses = np.stack([P00, P01, P10, P11], axis=1).astype(np.float64)
    buf["pose"] = poses
    buf["mesh"] = 0 if mesh_pairs is None else np.asarray(mesh_pairs, dtype=np.int32)
    if mesh_pairs is None:
        buf["mesh"][:, 1] = 1
    buf["seed_tri"] = 0 if seeds is None else np.asarray(seeds, dtype=np.int32)
    return qs


def ccd_batch(meshes, queries, params=None, stream=None):
    """Runs a query batch on the GPU; returns a structured array (RESULT_DTYPE)."""
    nq = len(queries)
    out = np.zeros(nq, dtype=RESULT_DTYPE)
    p = params if params is not None else default_params()
    _check(lib().c2a_ccd_batch(_mesh_table(meshes), len(meshes), queries, nq, C.byref(p),
                               out.ctypes.data_as(C.c_void_p), stream))
    return out


class Batch:
    """Device-resident batch: upload once, run many times (benchmarking)."""

    def __init__(self, meshes, queries, params=None, scorer=None):
        self._meshes = meshes
        self._scorer = scorer   # keeps a learned-order scorer alive with the batch
        self.nq = len(queries)
        p = params if params is not None else default_params()
        h = C.c_void_p()
        _check(lib().c2a_batch_create(_mesh_table(meshes), len(meshes), queries, self.nq,
                                      C.byref(p), C.byref(h)))
        self.handle = h

    def run(self, stream=None):
        _check(lib().c2a_batch_run(self.handle, stream))

    def run_graph(self, stream=None):
        """The same launch through a captured HIP graph (captured on first use)."""
        _check(lib().c2a_batch_run_graph(self.handle, stream))

    def fetch(self, stream=None):
        out = np.zeros(self.nq, dtype=RESULT_DTYPE)
        _check(lib().c2a_batch_fetch(self.handle, out.ctypes.data_as(C.c_void_p), stream))
        return out

    def last_ms(self):
        ms = C.c_float()
        _check(lib().c2a_batch_last_ms(self.handle, C.byref(ms)))
        return ms.value

    def tail_count(self):
        """Queries the last run handed to the tail pass (0: one launch)."""
        n = C.c_int64()
        _check(lib().c2a_batch_tail_count(self.handle, C.byref(n)))
        return n.value

    def samples(self, cap):
        """Training samples of the last run: (features (n, 32) float32, labels (n,),
        keys (n,) int64, selected count)."""
        x = np.zeros((cap, NF), dtype=np.float32)
        y = np.zeros(cap, dtype=np.float32)
        k = np.zeros(cap, dtype=np.int64)
        n, w = C.c_int64(), C.c_int64()
        _check(lib().c2a_batch_samples(self.handle, x.ctypes.data_as(C.c_void_p),
                                       y.ctypes.data_as(C.c_void_p), k.ctypes.data_as(C.c_void_p),
                                       C.c_int64(cap), C.byref(n), C.byref(w)))
        return x[:w.value], y[:w.value], k[:w.value], n.value

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().c2a_batch_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


def ccd_step(meshes, steps):
    nq = len(steps)
    out = (StepResult * nq)()
    _check(lib().c2a_ccd_step(_mesh_table(meshes), len(meshes), steps, nq, out, None))
    return out


def contact_count(meshes, poses, mesh_pairs, thresholds):
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 24)
    mp = np.ascontiguousarray(mesh_pairs, dtype=np.int32).reshape(-1, 2)
    th = np.ascontiguousarray(thresholds, dtype=np.float64)
    out = np.zeros(len(th), dtype=np.int32)
    _check(lib().c2a_contact_count(_mesh_table(meshes), len(meshes), _dp(poses), _ip(mp),
                                   _dp(th), len(th), _ip(out), None))
    return out


def contact_list(meshes, poses, mesh_pairs, thresholds, cap):
    """Contact records (C2A_QueryContact's cont_l, in visiting order): returns
    (counts (nq,), records (nq, cap) CONTACT_DTYPE); query i's first
    min(counts[i], cap) records are valid."""
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 24)
    mp = np.ascontiguousarray(mesh_pairs, dtype=np.int32).reshape(-1, 2)
    th = np.ascontiguousarray(thresholds, dtype=np.float64)
    counts = np.zeros(len(th), dtype=np.int32)
    recs = np.zeros((len(th), max(cap, 1)), dtype=CONTACT_DTYPE)
    _check(lib().c2a_contact_list(_mesh_table(meshes), len(meshes), _dp(poses), _ip(mp), _dp(th),
                                  len(th), int(cap), recs.ctypes.data_as(C.c_void_p), _ip(counts),
                                  None))
    return counts, recs[:, :cap]


def distance(meshes, poses, mesh_pairs, seed_tris=None, rel_err=0.0, abs_err=0.0,
             overlap_only=False):
    """Static distance queries (C2A_Distance): poses (nq, 24) = pose of mesh 1
    then mesh 2 (R row-major, T); returns a DISTANCE_DTYPE array.
    overlap_only=True: C2A_Collide's distance form (only pairs whose RSS
    volumes overlap are visited)."""
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 24)
    nq = len(poses)
    mp = np.ascontiguousarray(mesh_pairs, dtype=np.int32).reshape(nq, 2)
    seeds = None
    if seed_tris is not None:
        seeds = np.ascontiguousarray(seed_tris, dtype=np.int32).reshape(nq, 2)
    out = np.zeros(nq, dtype=DISTANCE_DTYPE)
    f = lib().c2a_collide_distance if overlap_only else lib().c2a_distance
    _check(f(_mesh_table(meshes), len(meshes), _dp(poses), _ip(mp),
             _ip(seeds) if seeds is not None else None, rel_err, abs_err, nq, out.ctypes.data,
             None))
    return out


def _static_args(poses, mesh_pairs):
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 24)
    nq = len(poses)
    return poses, nq, np.ascontiguousarray(mesh_pairs, dtype=np.int32).reshape(nq, 2)


def _per_query(v, nq):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (nq,)))


def distance_queue(meshes, poses, mesh_pairs, qsize, seed_tris=None, rel_err=0.0,
                   abs_err=0.0):
    """C2A_Distance with a queue of qsize tests (> 2: best-first search);
    returns a DISTANCE_DTYPE array."""
    poses, nq, mp = _static_args(poses, mesh_pairs)
    seeds = None
    if seed_tris is not None:
        seeds = np.ascontiguousarray(seed_tris, dtype=np.int32).reshape(nq, 2)
    out = np.zeros(nq, dtype=DISTANCE_DTYPE)
    _check(lib().c2a_distance_queue(_mesh_table(meshes), len(meshes), _dp(poses), _ip(mp),
                                    _ip(seeds) if seeds is not None else None, rel_err, abs_err,
                                    int(qsize), nq, out.ctypes.data, None))
    return out


def distance_contact_space(meshes, poses, mesh_pairs, upper_bounds, seed_tris=None):
    """C2A_Distance_At_ContactSpace; returns a STATIC_DTYPE array."""
    poses, nq, mp = _static_args(poses, mesh_pairs)
    up = _per_query(upper_bounds, nq)
    seeds = None
    if seed_tris is not None:
        seeds = np.ascontiguousarray(seed_tris, dtype=np.int32).reshape(nq, 2)
    out = np.zeros(nq, dtype=STATIC_DTYPE)
    _check(lib().c2a_distance_contact_space(_mesh_table(meshes), len(meshes), _dp(poses),
                                            _ip(mp), _ip(seeds) if seeds is not None else None,
                                            _dp(up), nq, out.ctypes.data, None))
    return out


def tolerance(meshes, poses, mesh_pairs, tolerances, qsize=2):
    """C2A_Tolerance; returns a STATIC_DTYPE array (found = closer)."""
    poses, nq, mp = _static_args(poses, mesh_pairs)
    tol = _per_query(tolerances, nq)
    out = np.zeros(nq, dtype=STATIC_DTYPE)
    _check(lib().c2a_tolerance(_mesh_table(meshes), len(meshes), _dp(poses), _ip(mp), _dp(tol),
                               int(qsize), nq, out.ctypes.data, None))
    return out


def collide_pairs(meshes, poses, mesh_pairs, flag=1, cap=1024):
    """C2A_Collide's PQP_CollideResult form; returns (STATIC_DTYPE array,
    pairs (nq, cap, 2) of build-order triangles, -1 past each count)."""
    poses, nq, mp = _static_args(poses, mesh_pairs)
    out = np.zeros(nq, dtype=STATIC_DTYPE)
    pairs = np.full((nq, cap, 2), -1, dtype=np.int32)
    _check(lib().c2a_collide_pairs(_mesh_table(meshes), len(meshes), _dp(poses), _ip(mp),
                                   int(flag), nq, int(cap), _ip(pairs), out.ctypes.data, None))
    return out, pairs


class Scorer:
    """The learned front-ordering pair scorer (GPU parameters)."""

    NPARAM = 6337

    def __init__(self, seed=0, path=None):
        h = C.c_void_p()
        if path is None:
            _check(lib().c2a_scorer_create(C.c_uint64(seed), C.byref(h)))
        else:
            _check(lib().c2a_scorer_load(os.fsencode(path), C.byref(h)))
        self.handle = h

    def save(self, path):
        _check(lib().c2a_scorer_save(self.handle, os.fsencode(path)))

    def __del__(self):
        if getattr(self, "handle", None):
            lib().c2a_scorer_destroy(self.handle)
            self.handle = None

    def params(self):
        p = np.zeros(self.NPARAM, dtype=np.float32)
        _check(lib().c2a_scorer_get_params(self.handle, p.ctypes.data_as(C.c_void_p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().c2a_scorer_set_params(self.handle, p.ctypes.data_as(C.c_void_p)))

    def infer(self, feats):
        x = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, 32)
        out = np.zeros(len(x), dtype=np.float32)
        _check(lib().c2a_scorer_infer(self.handle, x.ctypes.data_as(C.c_void_p), C.c_int64(len(x)),
                                      out.ctypes.data_as(C.c_void_p)))
        return out

    def grad(self, feats, labels):
        x = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, 32)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        g = np.zeros(self.NPARAM, dtype=np.float32)
        loss = C.c_float()
        _check(lib().c2a_scorer_grad(self.handle, x.ctypes.data_as(C.c_void_p),
                                     y.ctypes.data_as(C.c_void_p), C.c_int64(len(y)),
                                     g.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return loss.value, g

    def train(self, feats, labels, epochs=10, batch=1024, lr=1e-3, seed=0):
        x = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, 32)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        _check(lib().c2a_scorer_train(self.handle, x.ctypes.data_as(C.c_void_p),
                                      y.ctypes.data_as(C.c_void_p), C.c_int64(len(y)),
                                      int(epochs), int(batch), C.c_float(lr), C.c_uint64(seed),
                                      losses.ctypes.data_as(C.c_void_p)))
        return losses


def ccd_batch_noalloc(meshes, motions, mesh_pairs, seed_tris, nq, params=None, out=None,
                      stream=None):
    """c2a_ccd_batch_noalloc: every array argument is a device address (int)
    the caller owns -- motions nq x 2 x 24 doubles (motion_records), mesh
    pairs and seed pairs nq x 2 int32 (seeds may be None), out nq results.
    Returns once the work is queued on `stream`."""
    p = params if params is not None else default_params()
    _check(lib().c2a_ccd_batch_noalloc(_mesh_table(meshes), len(meshes), motions, mesh_pairs,
                                       seed_tris, int(nq), C.byref(p), out, stream))


def motion_records(P00, P01, P10, P11):
    """(nq, 2, 24) motion records (c2a_motion_prepare) of both objects."""
    nq = len(P00)
    out = np.zeros((nq, 2, 24))
    for i in range(nq):
        out[i, 0] = motion_prepare(P00[i], P01[i])
        out[i, 1] = motion_prepare(P10[i], P11[i])
    return out


def motion_prepare(pose_start, pose_end):
    a = np.ascontiguousarray(pose_start, dtype=np.float64)
    b = np.ascontiguousarray(pose_end, dtype=np.float64)
    m = np.zeros(24)
    _check(lib().c2a_motion_prepare(_dp(a), _dp(b), _dp(m)))
    return m


def motion_pose(motion24, t):
    m = np.ascontiguousarray(motion24, dtype=np.float64)
    pose, q = np.zeros(12), np.zeros(4)
    _check(lib().c2a_motion_pose(_dp(m), float(t), _dp(pose), _dp(q)))
    return pose, q
