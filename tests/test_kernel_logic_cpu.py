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
