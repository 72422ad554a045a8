"""ctypes binding of the FP64 CPU oracle (oracle/libc2a_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker.  The product library
never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libc2a_oracle.so")

d3 = C.c_double * 3
m33 = (C.c_double * 3) * 3


class Tri(C.Structure):
    _fields_ = [("p", m33), ("id", C.c_int), ("idx", C.c_int * 3)]


class BV(C.Structure):
    _fields_ = [("R", m33), ("Tr", d3), ("l", C.c_double * 2), ("r", C.c_double),
                ("R_loc", m33), ("angR", C.c_double), ("size", C.c_double),
                ("first_child", C.c_int), ("depth", C.c_int)]


class Model(C.Structure):
    _fields_ = [("ntri", C.c_int), ("tris", C.POINTER(Tri)), ("nbv", C.c_int),
                ("bvs", C.POINTER(BV)), ("last_tri", C.c_int)]


class Motion(C.Structure):
    _fields_ = [("R0", m33), ("T0", d3), ("R1", m33), ("T1", d3), ("q0", C.c_double * 4),
                ("cv", d3), ("axis", d3), ("angVel", C.c_double), ("axw", d3)]


class Result(C.Structure):
    _fields_ = [("toc", C.c_double), ("distance", C.c_double), ("mint", C.c_double),
                ("p1", d3), ("p2", d3), ("R_toc", m33), ("T_toc", d3),
                ("abs_err", C.c_double), ("rel_err", C.c_double), ("UpboundTOC", C.c_double),
                ("numCA", C.c_int), ("collisionfree", C.c_int), ("num_contact", C.c_int),
                ("num_bv_tests", C.c_longlong), ("num_tri_tests", C.c_longlong),
                ("last_triA", C.c_int), ("last_triB", C.c_int),
                ("out_last_triA", C.c_int), ("out_last_triB", C.c_int), ("flags", C.c_int),
                ("R", m33), ("T", d3)]


class Contact(C.Structure):
    _fields_ = [("triA", C.c_int), ("triB", C.c_int), ("typeA", C.c_int), ("typeB", C.c_int),
                ("fidA", C.c_int * 3), ("fidB", C.c_int * 3), ("PA", d3), ("PB", d3),
                ("distance", C.c_double)]


FLAG_ITER_CAP = 1
FLAG_TRANSLATION_ONLY = 2

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.c2o_rect_dist.restype = C.c_double
        L.c2o_rect_dist.argtypes = [dp, dp, dp, dp, dp, dp, dp]
        L.c2o_tri_dist.restype = C.c_double
        L.c2o_tri_dist.argtypes = [dp, dp, dp, dp, ip, ip, ip, ip]
        L.c2o_quat_from_matrix.argtypes = [dp, dp]
        L.c2o_matrix_from_quat.argtypes = [dp, dp]
        L.c2o_motion_init.argtypes = [C.POINTER(Motion), dp, dp, dp, dp]
        L.c2o_motion_integrate.argtypes = [C.POINTER(Motion), C.c_double, dp, dp, dp]
        L.c2o_motion_bound.restype = C.c_double
        L.c2o_motion_bound.argtypes = [C.POINTER(Motion), C.c_double, dp]
        L.c2o_motion_bound_leaf.restype = C.c_double
        L.c2o_motion_bound_leaf.argtypes = [C.POINTER(Motion), C.c_double, dp]
        L.c2o_sincos.argtypes = [C.c_double, dp, dp]
        L.c2o_set_libm_trig.argtypes = [C.c_int]
        L.c2o_set_solve_translation.argtypes = [C.c_int]
        L.c2o_set_strict.argtypes = [C.c_int]
        L.c2o_ca_loop_scripted.restype = C.c_int
        L.c2o_ca_loop_scripted.argtypes = [dp, dp, C.c_int, C.c_double, C.c_double, C.c_int,
                                           C.POINTER(Result), dp, dp, ip, C.c_int]
        L.c2o_get_strict.restype = C.c_int
        L.c2o_model_build.restype = C.c_int
        L.c2o_model_build.argtypes = [C.POINTER(Model), C.c_int, dp, ip, ip]
        L.c2o_model_free.argtypes = [C.POINTER(Model)]
        L.c2o_jacobi_eigen.argtypes = [dp, dp, dp]
        L.c2o_query_toc.restype = C.c_double
        L.c2o_query_toc.argtypes = [C.POINTER(Model), C.POINTER(Model), C.POINTER(Motion),
                                    C.POINTER(Motion), C.POINTER(Result), C.c_double,
                                    C.c_double, C.c_int]
        L.c2o_query_contact.restype = C.c_int
        L.c2o_query_contact.argtypes = [C.POINTER(Model), C.POINTER(Model), dp, dp, dp, dp,
                                        C.c_double, C.POINTER(Contact), C.c_int,
                                        C.POINTER(C.c_longlong)]
        L.c2o_distance.restype = None
        L.c2o_distance.argtypes = [C.POINTER(Model), C.POINTER(Model), dp, dp, dp, dp,
                                   C.c_double, C.c_double, C.POINTER(DistanceResult)]
        L.c2o_collide_distance.restype = None
        L.c2o_collide_distance.argtypes = L.c2o_distance.argtypes
        L.c2o_distance_batch.restype = None
        L.c2o_distance_batch.argtypes = [C.POINTER(Model), C.POINTER(Model), C.c_int, dp, ip,
                                         C.c_double, C.c_double, C.c_int,
                                         C.POINTER(DistanceResult)]
        L.c2o_toc_step.restype = None
        L.c2o_toc_step.argtypes = [C.POINTER(Model), C.POINTER(Model), C.POINTER(Motion),
                                   C.POINTER(Motion), dp, dp, dp, dp, C.POINTER(Model),
                                   C.POINTER(Model), C.POINTER(Result)]
        L.c2o_solve.restype = C.c_int
        L.c2o_solve.argtypes = [dp, dp, dp, dp, C.POINTER(Model), dp, dp, dp, dp,
                                C.POINTER(Model), dp, dp, dp, dp, C.POINTER(Result)]
        L.c2o_batch_toc.restype = C.c_int
        L.c2o_batch_toc.argtypes = [C.POINTER(Model), C.POINTER(Model), C.c_int, dp, dp, dp,
                                    dp, ip, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.POINTER(Result)]
        L.c2o_distance_queue_batch.restype = None
        L.c2o_distance_queue_batch.argtypes = [C.POINTER(Model), C.POINTER(Model), C.c_int, dp,
                                               ip, C.c_double, C.c_double, C.c_int,
                                               C.POINTER(DistanceResult)]
        L.c2o_distance_cs_batch.restype = None
        L.c2o_distance_cs_batch.argtypes = [C.POINTER(Model), C.POINTER(Model), C.c_int, dp, ip,
                                            dp, C.POINTER(DistanceResult), ip]
        L.c2o_tolerance_batch.restype = None
        L.c2o_tolerance_batch.argtypes = [C.POINTER(Model), C.POINTER(Model), C.c_int, dp, dp,
                                          C.c_int, C.POINTER(ToleranceResult)]
        L.c2o_collide_pairs_batch.restype = None
        L.c2o_collide_pairs_batch.argtypes = [C.POINTER(Model), C.POINTER(Model), C.c_int, dp,
                                              C.c_int, C.c_int, ip, C.POINTER(CollideResult)]
        L.c2o_tri_contact.restype = C.c_int
        L.c2o_tri_contact.argtypes = [dp, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def rect_dist(Rab, Tab, a, b):
    Rab, Tab, a, b = f64(Rab), f64(Tab), f64(a), f64(b)
    P, Q, S = np.zeros(3), np.full(3, np.nan), np.full(3, np.nan)
    d = lib().c2o_rect_dist(_dp(Rab), _dp(Tab), _dp(a), _dp(b), _dp(P), _dp(Q), _dp(S))
    return d, P, Q, S


def tri_dist(S, T):
    S, T = f64(S), f64(T)
    P, Q = np.zeros(3), np.zeros(3)
    ta, fa, tb, fb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    d = lib().c2o_tri_dist(_dp(P), _dp(Q), _dp(S), _dp(T), C.byref(ta), C.byref(fa),
                           C.byref(tb), C.byref(fb))
    return d, P, Q, (ta.value, fa.value, tb.value, fb.value)


def quat_from_matrix(R):
    R = f64(R)
    q = np.zeros(4)
    lib().c2o_quat_from_matrix(_dp(R), _dp(q))
    return q


def matrix_from_quat(q):
    q = f64(q)
    R = np.zeros((3, 3))
    lib().c2o_matrix_from_quat(_dp(q), _dp(R))
    return R


def sincos(x):
    s, c = C.c_double(), C.c_double()
    lib().c2o_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def set_libm_trig(flag):
    lib().c2o_set_libm_trig(int(flag))


def set_solve_translation(flag):
    """C2A_Solve's translation-only branch for queries whose two rotations are
    below 1e-8 (default on)."""
    lib().c2o_set_solve_translation(int(flag))


def ca_loop_scripted(dist, mint, tol_d, tol_t, max_iter=150, cap=512):
    """The CA loop with a scripted step (call k returns dist[k], mint[k]).
    Returns (Result, calls, lamdas, upbounds, numCAs) of the step calls."""
    dist, mint = f64(dist), f64(mint)
    assert len(dist) == len(mint) > 0
    res = Result()
    lam, up = np.zeros(cap), np.zeros(cap)
    nca = np.zeros(cap, dtype=np.int32)
    n = lib().c2o_ca_loop_scripted(_dp(dist), _dp(mint), len(dist), float(tol_d), float(tol_t),
                                   int(max_iter), C.byref(res), _dp(lam), _dp(up), _ip(nca),
                                   cap)
    k = min(n, cap)
    return res, n, lam[:k], up[:k], nca[:k]


def set_strict(flag):
    """Strict (conservative) CA mode: world-frame axis in the motion bounds,
    every tested leaf pair bounds mint, touching leaves bound it by 0.
    Off (default) = the reference's semantics."""
    lib().c2o_set_strict(int(flag))


class strict_mode:
    """Context manager: the oracle in strict mode inside the block."""

    def __init__(self, on=True):
        self.on = on

    def __enter__(self):
        self.prev = lib().c2o_get_strict()
        set_strict(self.on)
        return self

    def __exit__(self, *exc):
        set_strict(self.prev)
        return False


def jacobi_eigen(A):
    A = f64(A)
    V, d = np.zeros((3, 3)), np.zeros(3)
    lib().c2o_jacobi_eigen(_dp(A), _dp(V), _dp(d))
    return V, d


class OMotion:
    def __init__(self, R0, T0, R1, T1):
        self.m = Motion()
        lib().c2o_motion_init(C.byref(self.m), _dp(f64(R0)), _dp(f64(T0)), _dp(f64(R1)),
                              _dp(f64(T1)))

    @property
    def angVel(self):
        return self.m.angVel

    @property
    def axis(self):
        return np.array(self.m.axis[:])

    @property
    def cv(self):
        return np.array(self.m.cv[:])

    def integrate(self, t):
        R, T, q7 = np.zeros((3, 3)), np.zeros(3), np.zeros(7)
        lib().c2o_motion_integrate(C.byref(self.m), float(t), _dp(R), _dp(T), _dp(q7))
        return R, T, q7

    def bound(self, angR, N):
        return lib().c2o_motion_bound(C.byref(self.m), float(angR), _dp(f64(N)))

    def bound_leaf(self, angR, S):
        return lib().c2o_motion_bound_leaf(C.byref(self.m), float(angR), _dp(f64(S)))


class OModel:
    """A model built by the oracle's PQP-style RSS builder."""

    def __init__(self, tris, ids=None, idx=None):
        tris = f64(tris).reshape(-1, 9)
        n = tris.shape[0]
        self.m = Model()
        idsa = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        idxa = None if idx is None else np.ascontiguousarray(idx, dtype=np.int32).reshape(-1)
        rc = lib().c2o_model_build(C.byref(self.m), n, _dp(tris),
                                   None if idsa is None else _ip(idsa),
                                   None if idxa is None else _ip(idxa))
        if rc != 0:
            raise ValueError("oracle model build failed: %d" % rc)

    def __del__(self):
        try:
            lib().c2o_model_free(C.byref(self.m))
        except Exception:
            pass

    @property
    def ntri(self):
        return self.m.ntri

    @property
    def nbv(self):
        return self.m.nbv

    @property
    def last_tri(self):
        return self.m.last_tri

    @last_tri.setter
    def last_tri(self, v):
        self.m.last_tri = int(v)

    def tris(self):
        buf = (Tri * self.m.ntri).from_address(C.addressof(self.m.tris.contents))
        a = np.ctypeslib.as_array(buf)
        return a

    def tri_xyz(self):
        """Reordered triangles as an (n, 3, 3) array."""
        return np.array(self.tris()["p"], dtype=np.float64)

    def bvs(self):
        return [self.m.bvs[i] for i in range(self.m.nbv)]

    def bv_arrays(self):
        """Node arrays as numpy (structured copy)."""
        n = self.m.nbv
        raw = np.ctypeslib.as_array((BV * n).from_address(C.addressof(self.m.bvs.contents)))
        return raw.copy()


def query_toc(m1, m2, pose00, pose01, pose10, pose11, tol_d, tol_t, seed=None,
              max_iter=150):
    """C2A_QueryTimeOfContact on one query.  pose = (R 3x3, T 3)."""
    mo1 = OMotion(pose00[0], pose00[1], pose01[0], pose01[1])
    mo2 = OMotion(pose10[0], pose10[1], pose11[0], pose11[1])
    res = Result()
    if seed is None:
        seed = (m1.last_tri, m2.last_tri)
    res.last_triA, res.last_triB = int(seed[0]), int(seed[1])
    lib().c2o_query_toc(C.byref(m1.m), C.byref(m2.m), C.byref(mo1.m), C.byref(mo2.m),
                        C.byref(res), float(tol_d), float(tol_t), int(max_iter))
    return res


def query_contact(m1, m2, R1, T1, R2, T2, threshold, cap=0):
    out = (Contact * max(cap, 1))()
    bvt = C.c_longlong()
    n = lib().c2o_query_contact(C.byref(m1.m), C.byref(m2.m), _dp(f64(R1)), _dp(f64(T1)),
                                _dp(f64(R2)), _dp(f64(T2)), float(threshold), out, int(cap),
                                C.byref(bvt))
    return n, [out[i] for i in range(min(n, cap))], bvt.value


class DistanceResult(C.Structure):
    _fields_ = [("distance", C.c_double), ("p1", C.c_double * 3), ("p2", C.c_double * 3),
                ("bv_tests", C.c_longlong), ("tri_tests", C.c_longlong),
                ("last_tri", C.c_int * 2)]


def distance(m1, m2, R1, T1, R2, T2, rel_err=0.0, abs_err=0.0, seeds=None,
             overlap_only=False):
    """C2A_Distance (qsize <= 2) at explicit poses, or with overlap_only
    C2A_Collide's distance form.  seeds sets the models' last_tri first
    (build order); both are updated like the reference's."""
    if seeds is not None:
        m1.last_tri, m2.last_tri = int(seeds[0]), int(seeds[1])
    r = DistanceResult()
    f = lib().c2o_collide_distance if overlap_only else lib().c2o_distance
    f(C.byref(m1.m), C.byref(m2.m), _dp(f64(R1)), _dp(f64(T1)), _dp(f64(R2)),
                       _dp(f64(T2)), float(rel_err), float(abs_err), C.byref(r))
    return r


def distance_batch(m1, m2, poses24, seeds=None, rel_err=0.0, abs_err=0.0, nthreads=1):
    """Independent distance queries; returns the list of DistanceResult."""
    P = f64(poses24).reshape(-1, 24)
    nq = len(P)
    out = (DistanceResult * nq)()
    sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.int32).reshape(nq, 2)
    lib().c2o_distance_batch(C.byref(m1.m), C.byref(m2.m), nq, _dp(P),
                             None if sd is None else _ip(sd), float(rel_err), float(abs_err),
                             int(nthreads), out)
    return list(out)


class ToleranceResult(C.Structure):
    _fields_ = [("closer", C.c_int), ("tri", C.c_int * 2), ("tolerance", C.c_double),
                ("distance", C.c_double), ("p1", C.c_double * 3), ("p2", C.c_double * 3),
                ("bv_tests", C.c_longlong), ("tri_tests", C.c_longlong)]


class CollideResult(C.Structure):
    _fields_ = [("bv_tests", C.c_longlong), ("tri_tests", C.c_longlong),
                ("num_pairs", C.c_int), ("pad", C.c_int)]


def _poses(poses24):
    P = f64(poses24).reshape(-1, 24)
    return P, len(P)


def _seeds(seeds, nq):
    return None if seeds is None else np.ascontiguousarray(seeds, dtype=np.int32).reshape(nq, 2)


def distance_queue_batch(m1, m2, poses24, qsize, seeds=None, rel_err=0.0, abs_err=0.0):
    """C2A_Distance with the given qsize (> 2: the best-first queue search)."""
    P, nq = _poses(poses24)
    out = (DistanceResult * nq)()
    sd = _seeds(seeds, nq)
    lib().c2o_distance_queue_batch(C.byref(m1.m), C.byref(m2.m), nq, _dp(P),
                                   None if sd is None else _ip(sd), float(rel_err),
                                   float(abs_err), int(qsize), out)
    return list(out)


def distance_cs_batch(m1, m2, poses24, upper, seeds=None):
    """C2A_Distance_At_ContactSpace; returns (results, found flags)."""
    P, nq = _poses(poses24)
    out = (DistanceResult * nq)()
    found = np.zeros(nq, np.int32)
    up = np.ascontiguousarray(np.broadcast_to(np.asarray(upper, np.float64), (nq,)))
    sd = _seeds(seeds, nq)
    lib().c2o_distance_cs_batch(C.byref(m1.m), C.byref(m2.m), nq, _dp(P),
                                None if sd is None else _ip(sd), _dp(up), out, _ip(found))
    return list(out), found


def tolerance_batch(m1, m2, poses24, tolerance, qsize=2):
    """C2A_Tolerance per query (tolerance: scalar or one per query)."""
    P, nq = _poses(poses24)
    out = (ToleranceResult * nq)()
    tol = np.ascontiguousarray(np.broadcast_to(np.asarray(tolerance, np.float64), (nq,)))
    lib().c2o_tolerance_batch(C.byref(m1.m), C.byref(m2.m), nq, _dp(P), _dp(tol), int(qsize),
                              out)
    return list(out)


def collide_pairs_batch(m1, m2, poses24, flag=1, cap=1024):
    """C2A_Collide's PQP_CollideResult form; returns (results, pairs (nq, cap, 2))."""
    P, nq = _poses(poses24)
    out = (CollideResult * nq)()
    pairs = np.full((nq, cap, 2), -1, np.int32)
    lib().c2o_collide_pairs_batch(C.byref(m1.m), C.byref(m2.m), nq, _dp(P), int(flag), int(cap),
                                  _ip(pairs), out)
    return list(out), pairs


def tri_contact(A, B):
    """PQP's TriContact on two (3, 3) vertex arrays in one frame."""
    A, B = f64(A).reshape(3, 3), f64(B).reshape(3, 3)
    return lib().c2o_tri_contact(_dp(A[0]), _dp(A[1]), _dp(A[2]), _dp(B[0]), _dp(B[1]),
                                 _dp(B[2]))


def toc_step(m1, m2, mo1, mo2, R1, T1, R2, T2, seeds, numCA, mint_prev, upbound):
    """C2A_TimeOfContactStep at explicit poses: seeds = res->last_triA/B; updates
    the models' last_tri; returns the Result."""
    res = Result()
    res.last_triA, res.last_triB = int(seeds[0]), int(seeds[1])
    res.numCA = int(numCA)
    res.mint = float(mint_prev)
    res.UpboundTOC = float(upbound)
    lib().c2o_toc_step(C.byref(m1.m), C.byref(m2.m), C.byref(mo1.m), C.byref(mo2.m),
                       _dp(f64(R1)), _dp(f64(T1)), _dp(f64(R2)), _dp(f64(T2)),
                       C.byref(m1.m), C.byref(m2.m), C.byref(res))
    return res


def solve(m1, m2, pose00, pose01, pose10, pose11):
    """C2A_Solve; seeds from and updates the models' last_tri."""
    res = Result()
    R0o, T0o, R1o, T1o = np.zeros((3, 3)), np.zeros(3), np.zeros((3, 3)), np.zeros(3)
    lib().c2o_solve(_dp(f64(pose00[0])), _dp(f64(pose00[1])), _dp(f64(pose01[0])),
                    _dp(f64(pose01[1])), C.byref(m1.m), _dp(f64(pose10[0])),
                    _dp(f64(pose10[1])), _dp(f64(pose11[0])), _dp(f64(pose11[1])),
                    C.byref(m2.m), _dp(R0o), _dp(T0o), _dp(R1o), _dp(T1o), C.byref(res))
    return res, (R0o, T0o), (R1o, T1o)


RESULT_FIELDS = ["toc", "distance", "numCA", "collisionfree", "num_contact",
                 "num_bv_tests", "num_tri_tests", "out_last_triA", "out_last_triB", "flags"]


def batch_toc(m1, m2, P00, P01, P10, P11, tol_d, tol_t, seeds=None, max_iter=150,
              with_contacts=False, chain=False, nthreads=1):
    """Run nq queries; P** are (nq, 12) arrays (row-major R then T).
    Returns a structured numpy array of Result records."""
    P00, P01, P10, P11 = f64(P00), f64(P01), f64(P10), f64(P11)
    nq = P00.shape[0]
    out = (Result * nq)()
    sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.int32)
    lib().c2o_batch_toc(C.byref(m1.m), C.byref(m2.m), nq, _dp(P00), _dp(P01), _dp(P10),
                        _dp(P11), None if sd is None else _ip(sd), float(tol_d), float(tol_t),
                        int(max_iter), int(with_contacts), int(chain), int(nthreads), out)
    return np.ctypeslib.as_array(out).copy()
