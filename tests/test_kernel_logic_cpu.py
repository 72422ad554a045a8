"""The GPU kernel's per-query logic (c2a_amd/csrc/ccd_core.h) compiled for the
host and run one lane at a time by tests/emu/ccd_emu.cpp, against the oracle.
This checks the walk / speculation logic on machines without a GPU; the GPU
tests check the same code on the device."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu", "libccd_emu.so")


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        import __graft_entry__
        __graft_entry__._build_emu()
    return C.CDLL(EMU)


def run_emu(emu, t1, t2, P, params, K, seeds=None, mesh_pairs=None):
    nq = len(P[0])
    qs = G.make_queries(*P, mesh_pairs=mesh_pairs, seeds=seeds)
    out = np.zeros(nq, dtype=G.RESULT_DTYPE)
    a = np.ascontiguousarray(t1.reshape(-1, 9))
    b = np.ascontiguousarray(t2.reshape(-1, 9))
    rc = emu.emu_ccd_batch(G._dp(a), C.c_int64(len(a)), G._dp(b), C.c_int64(len(b)), qs,
                           C.c_int64(nq), C.byref(params), out.ctypes.data_as(C.c_void_p), K)
    assert rc == 0
    return out


def same(ref, got):
    for f, g in [("toc", "toc"), ("distance", "distance"), ("mint", "mint"), ("numCA", "numCA"),
                 ("collisionfree", "collisionfree"), ("num_bv_tests", "bv_tests"),
                 ("num_tri_tests", "tri_tests"), ("num_contact", "num_contact")]:
        assert np.array_equal(ref[f], got[g]), f
    assert np.array_equal(ref["out_last_triA"], got["last_tri"][:, 0])
    assert np.array_equal(ref["out_last_triB"], got["last_tri"][:, 1])


@pytest.mark.parametrize("K", [1, 2, 3, 4, 5, 6])
def test_kernel_logic_knot(emu, K):
    tris, _ = W.torus_knot(n_seg=128, n_ring=12)
    t9 = tris.reshape(-1, 9)
    nq = 24
    P = W.knot_queries(nq, 500 + K, collide_frac=0.4)
    got = run_emu(emu, t9, t9, P, G.default_params(tol_d=1e-3, tol_t=1e-4), K)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                      with_contacts=True, nthreads=4)
    same(ref, got)


@pytest.mark.parametrize("ext,adaptive", [(1, 0), (3, 0), (1, 1), (2, 1)])
def test_kernel_logic_replay_extension(emu, ext, adaptive):
    """Pairs computed ahead below replayed pairs (ext_levels) change no result,
    at every step or switched per step by the previous step's misses (the
    library's mode for the autotuned levels)."""
    tris, _ = W.torus_knot(n_seg=128, n_ring=12)
    t9 = tris.reshape(-1, 9)
    nq = 24
    P = W.knot_queries(nq, 610 + ext, collide_frac=0.4)
    emu.emu_set_ext_adaptive.argtypes = [C.c_int]
    emu.emu_set_ext_adaptive(adaptive)
    try:
        got = run_emu(emu, t9, t9, P, G.default_params(tol_d=1e-3, tol_t=1e-4, ext_levels=ext),
                      5)
    finally:
        emu.emu_set_ext_adaptive(0)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                      with_contacts=True, nthreads=4)
    same(ref, got)


def test_kernel_logic_unequal_meshes_and_seeds(emu):
    a, _ = W.icosphere(3, 0.1, 21)
    b, _ = W.torus(24, 12)
    a, b = a.reshape(-1, 9) * 0.8, b.reshape(-1, 9)
    nq = 24
    P = W.knot_queries(nq, 77, collide_frac=0.5)
    seeds = np.stack([np.arange(nq) * 13 % len(a), np.arange(nq) * 7 % len(b)], 1).astype(np.int32)
    got = run_emu(emu, a, b, P, G.default_params(), 5, seeds=seeds)
    m1, m2 = O.OModel(a), O.OModel(b)
    ref = O.batch_toc(m1, m2, *P, 1e-4, 1e-4, seeds=seeds, with_contacts=True, nthreads=4)
    same(ref, got)


@pytest.fixture()
def emu_learned(emu):
    emu.emu_set_scorer.argtypes = [C.c_void_p]
    emu.emu_set_sampling.argtypes = [C.c_double, C.c_int64]
    emu.emu_samples.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    emu.emu_samples.restype = C.c_int64
    yield emu
    emu.emu_set_scorer(None)
    emu.emu_set_sampling(0.0, 0)


def emu_samples(emu, cap):
    rows = np.zeros((cap, G.NF + 1), np.float32)
    keys = np.zeros(cap, np.int64)
    n = emu.emu_samples(rows.ctypes.data, keys.ctypes.data, cap)
    w = min(n, cap)
    o = np.argsort(keys[:w], kind="stable")
    return rows[:w][o], keys[:w][o], n


def _knot_case(nq=16, seed=900):
    tris, _ = W.torus_knot(n_seg=128, n_ring=12)
    t9 = tris.reshape(-1, 9)
    return t9, W.knot_queries(nq, seed, collide_frac=0.4)


def test_sampling_keeps_results_and_selects_by_hash(emu_learned):
    """Logging training samples changes no result; the hash selection at a
    lower rate picks a subset of the same rows."""
    emu = emu_learned
    t9, P = _knot_case()
    params = G.default_params(tol_d=1e-3, tol_t=1e-4)
    base = run_emu(emu, t9, t9, P, params, 5)
    emu.emu_set_sampling(1.0, 400000)
    full = run_emu(emu, t9, t9, P, params, 5)
    rows1, keys1, n1 = emu_samples(emu, 400000)
    emu.emu_set_sampling(0.25, 400000)
    run_emu(emu, t9, t9, P, params, 5)
    rows2, keys2, n2 = emu_samples(emu, 400000)
    for f in base.dtype.names:
        assert np.array_equal(base[f], full[f]), f
    assert n1 == len(keys1) > 1000 and len(np.unique(keys1)) == n1
    assert 0.15 * n1 < n2 < 0.35 * n1
    pos = np.searchsorted(keys1, keys2)
    assert np.array_equal(keys1[pos], keys2)
    assert np.array_equal(rows1[pos], rows2)
    # every sample is a replayed pair below the root of a step after the first
    step = (keys1 >> 26) & 0xff
    rec = keys1 & 0x3ffffff
    assert (rec > 0).all() and (step >= 1).all()
    assert np.isfinite(rows1).all()
    # leaf-flag feature: (leaf1 + 2 leaf2) / 3; a leaf pair's label is its own
    # triangle distance, so it is never below log(1e-4)
    lf = np.round(rows1[:, 31] * 3).astype(int)
    assert set(np.unique(lf)) <= {0, 1, 2, 3} and (lf == 3).any()
    assert (rows1[:, G.NF] >= np.float32(np.log(1e-4)) - 1e-6).all()


def test_learned_order_is_conservative_on_cpu(emu_learned):
    """A (random) scorer only reorders siblings: every query still ends with
    the plain order's collision flag, no colliding query ends with its TOC
    pose in contact unless the reference order's does too, and the
    traversal differs."""
    emu = emu_learned
    t9, P = _knot_case(24, 901)
    params = G.default_params(tol_d=1e-3, tol_t=1e-4)
    plain = run_emu(emu, t9, t9, P, params, 5)
    rng = np.random.default_rng(5)
    p = (rng.standard_normal(6337) * 0.2).astype(np.float32)
    emu.emu_set_scorer(p.ctypes.data)
    learned = run_emu(emu, t9, t9, P, params, 5)
    assert np.array_equal(plain["collisionfree"], learned["collisionfree"])

    def over(r):
        return (r["collisionfree"] == 0) & (r["toc"] > 0) & (r["distance"] == 0)
    assert not (over(learned) & ~over(plain)).any()
    assert not np.array_equal(plain["bv_tests"], learned["bv_tests"])


def factorial_chain(n, seed, mirror=False):
    """Triangles along x at k!/(n-1)!, k = 0..n-1, each scaled with its
    position: every mean split peels off the farthest one, so the hierarchy
    is a chain about 0.7 n levels deep (the tiniest ones merge).  mirror puts
    it on the negative x axis."""
    import math
    rng = np.random.default_rng(seed)
    top = float(math.factorial(n - 1))
    out = []
    for i in range(n):
        x = math.factorial(i) / top
        t = np.array([[0, 0, 0], [0.3, 0, 0], [0, 0.3, 0.1]]) * x * (1 + 0.1 * rng.random())
        t = t + np.array([x, 0.2 * x * rng.random(), 0.2 * x * rng.random()])
        if mirror:
            t[:, 0] *= -1
        out.append(t.ravel())
    return np.array(out)


def chain_queries(nq, seed):
    """The mirrored chain moving from a small gap across the first chain's
    tip while turning slightly about z."""
    rng = np.random.default_rng(seed)
    P = [np.zeros((nq, 12)) for _ in range(4)]
    for i in range(nq):
        for k in range(4):
            P[k][i, :9] = np.eye(3).ravel()
        a = rng.uniform(1e-3, 3e-2)
        c, s = np.cos(a), np.sin(a)
        P[3][i, :9] = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]).ravel()
        g = rng.uniform(1e-3, 5e-3)
        P[2][i, 9:] = [-g, rng.uniform(-1e-3, 1e-3), 0]
        P[3][i, 9:] = [2 * g, P[2][i, 10], 0]
    return P


def test_kernel_logic_deep_hierarchies(emu):
    """Two hierarchies about 70 levels deep: the pair runs on the 192-entry
    stack (the walks themselves stay shallow: the error bounds prune the
    chains' tips, which all lie within the gap)."""
    t1, t2 = factorial_chain(100, 1), factorial_chain(100, 2, mirror=True)
    assert G.Mesh(t1).max_depth + G.Mesh(t2).max_depth + 2 > 96
    nq = 12
    P = chain_queries(nq, 808)
    seeds = np.full((nq, 2), 99, np.int32)
    m1, m2 = O.OModel(t1), O.OModel(t2)
    ref = O.batch_toc(m1, m2, *P, 1e-6, 1e-6, seeds=seeds, with_contacts=True)
    got = run_emu(emu, t1, t2, P, G.default_params(tol_d=1e-6, tol_t=1e-6), 5, seeds=seeds)
    same(ref, got)


@pytest.mark.parametrize("lanes", [1, 64])
@pytest.mark.parametrize("budget", [50, 3000])
def test_kernel_logic_tail_hand_over(emu, lanes, budget):
    """The tail pass's hand-over: a first run hands every query that has made
    `budget` BV tests at the end of a CA step over (saving its CA-loop
    state), a second run continues them without a replay for their first
    step.  Results equal the oracle's, with one lane and with 64 (the
    second run here also has one wavefront; the kernel's tail pass has
    four, whose code is the same)."""
    lib = emu if lanes == 1 else C.CDLL(os.path.join(os.path.dirname(EMU), "libccd_emu64.so"))
    lib.emu_set_tail_budget.argtypes = [C.c_int64]
    lib.emu_tail_count.restype = C.c_int64
    tris, _ = W.torus_knot(n_seg=64, n_ring=8)
    t9 = tris.reshape(-1, 9)
    nq = 16
    P = W.knot_queries(nq, 123, 0.4)
    ref = O.batch_toc(O.OModel(t9), O.OModel(t9), *P, 1e-3, 1e-4,
                      seeds=np.zeros((nq, 2), np.int32), with_contacts=True)
    try:
        lib.emu_set_tail_budget(budget)
        for strict in (0, 1):
            r = ref
            if strict:
                with O.strict_mode():
                    r = O.batch_toc(O.OModel(t9), O.OModel(t9), *P, 1e-3, 1e-4,
                                    seeds=np.zeros((nq, 2), np.int32), with_contacts=True)
            got = run_emu(lib, t9, t9, P, G.default_params(tol_d=1e-3, tol_t=1e-4, strict=strict),
                          5)
            same(r, got)
            assert 0 < lib.emu_tail_count() <= int((r["num_bv_tests"] >= budget).sum())
    finally:
        lib.emu_set_tail_budget(0)
