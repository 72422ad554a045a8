r"""The CCD kernel's per-wavefront code with 64 lanes on the CPU.

tests/emu/libccd_emu64.so runs ccd_core.h's walk, replay, scan and rounds
with 64 lanes on fibers (tests/emu/simt.cpp): every wavefront primitive is
a barrier that exchanges the lanes' values, and between barriers the lanes
run one after another instead of in lockstep.  So the lane-strided work of
a round (each lane's slots), the cross-lane hand-offs through LDS and the
arena, and the wave-wide scan decisions all run as on the GPU, and a
missing barrier shows up as a wrong result.  Lanes that reach different
primitives, or a lane that finishes while others wait, fail the run.

Each case is compared field by field with the FP64 oracle, including the
arena-backed rounds (KArgs.arena_rounds) that the library keeps switched
off, small replay logs (the log-full paths), the replay's extension, the
strict mode, the learned order, and the rounds that several misses of a
scan window share (one lane decides one event per scan pass, so only the
64-lane run gathers them).

The same tests run under AddressSanitizer by pointing CCD_EMU64 at an ASan
build of the same sources:

    g++ -O1 -g -std=c++17 -ffp-contract=off -fPIC -shared -fsanitize=address \
        -DC2A_EMU_SIMT -o /tmp/libccd_emu64_asan.so tests/emu/ccd_emu.cpp \
        tests/emu/simt.cpp c2a_amd/csrc/bvh.cpp
    CCD_EMU64=/tmp/libccd_emu64_asan.so ASAN_OPTIONS=detect_leaks=0 \
        LD_PRELOAD=$(gcc -print-file-name=libasan.so) \
        python -m pytest tests/test_kernel_simt_cpu.py"""
import ctypes as C
import os

import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O
from tests.test_kernel_logic_cpu import run_emu, same

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def emu64():
    L = C.CDLL(os.environ.get("CCD_EMU64", os.path.join(HERE, "emu", "libccd_emu64.so")))
    L.emu_set_arena_rounds.argtypes = [C.c_int]
    L.emu_set_scorer.argtypes = [C.c_void_p]
    L.emu_set_ext_adaptive.argtypes = [C.c_int]
    yield L
    L.emu_set_arena_rounds(1)
    L.emu_set_scorer(None)


@pytest.fixture(scope="module")
def knot():
    tris, _ = W.torus_knot(n_seg=64, n_ring=8)
    t9 = tris.reshape(-1, 9)
    nq = 10
    P = W.knot_queries(nq, 123, 0.4)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                      with_contacts=True)
    return t9, P, ref


CASES = [
    dict(K=5, arena=0),
    dict(K=5, arena=1),
    dict(K=1, arena=1),
    dict(K=3, arena=1, miss_depth=1),
    dict(K=6, arena=1),
    dict(K=5, arena=1, replay_log=64),      # the log fills: records are dropped
    dict(K=5, arena=0, replay_log=256),
    dict(K=5, arena=1, replay_log=0),       # no replay at all
    dict(K=5, arena=1, ext_levels=1),
    dict(K=4, arena=0, ext_levels=2),
    dict(K=5, arena=0, ext_levels=1, adaptive=1),   # extension switched per step
    dict(K=5, arena=0, miss_depth=1),       # 16 misses per shared round
    dict(K=5, arena=0, miss_depth=3),       # 4 per round
    dict(K=4, arena=0, miss_depth=3),       # 2 per round
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join("%s%s" % kv for kv in c.items()))
def test_simt_bit_exact(emu64, knot, case):
    t9, P, ref = knot
    c = dict(case)
    K, arena, adaptive = c.pop("K"), c.pop("arena"), c.pop("adaptive", 0)
    emu64.emu_set_arena_rounds(arena)
    emu64.emu_set_ext_adaptive(adaptive)
    try:
        got = run_emu(emu64, t9, t9, P,
                      G.default_params(tol_d=1e-3, tol_t=1e-4, spec_depth=K, **c), K)
    finally:
        emu64.emu_set_ext_adaptive(0)
    same(ref, got)


def test_simt_strict(emu64, knot):
    t9, P, _ = knot
    nq = len(P[0])
    m1, m2 = O.OModel(t9), O.OModel(t9)
    with O.strict_mode():
        ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                          with_contacts=True)
    emu64.emu_set_arena_rounds(1)
    got = run_emu(emu64, t9, t9, P, G.default_params(tol_d=1e-3, tol_t=1e-4, strict=1), 5)
    same(ref, got)


def test_simt_learned_order_matches_one_lane(emu64, knot):
    """Under the learned order the 64-lane run and the one-lane run visit the
    same pairs (same scores, same decisions)."""
    t9, P, _ = knot
    emu1 = C.CDLL(os.path.join(HERE, "emu", "libccd_emu.so"))
    emu1.emu_set_scorer.argtypes = [C.c_void_p]
    rng = np.random.default_rng(3)
    params = (rng.normal(size=G.Scorer.NPARAM) * 0.1).astype(np.float32)
    outs = []
    for lib in (emu1, emu64):
        lib.emu_set_scorer(params.ctypes.data)
        try:
            outs.append(run_emu(lib, t9, t9, P,
                                G.default_params(tol_d=1e-3, tol_t=1e-4, spec_depth=5), 5))
        finally:
            lib.emu_set_scorer(None)
    for f in ("toc", "distance", "mint", "numCA", "bv_tests", "tri_tests", "last_tri"):
        assert np.array_equal(outs[0][f], outs[1][f]), f


@pytest.mark.parametrize("miss_depth", [1, 2, 3])
def test_simt_misses_share_rounds(emu64, knot, miss_depth):
    """Misses found in one 64-event scan window share a round (a shallow heap
    below each): bit-exact, and fewer rounds than the one-lane run, whose
    scan sees one event at a time."""
    t9, P, ref = knot
    emu1 = C.CDLL(os.path.join(HERE, "emu", "libccd_emu.so"))
    rounds = []
    for lib in (emu1, emu64):
        lib.emu_rounds.restype = C.c_int64
        lib.emu_set_arena_rounds(0)
        got = run_emu(lib, t9, t9, P, G.default_params(tol_d=1e-3, tol_t=1e-4, spec_depth=5,
                                                        miss_depth=miss_depth), 5)
        same(ref, got)
        rounds.append(lib.emu_rounds())
    assert rounds[1] < 0.95 * rounds[0], rounds
