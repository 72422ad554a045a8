"""Learned front ordering on the GPU: the training samples the exact GPU
traversal logs, the scorer inside the traversal (fewer BV tests, no more
overshoots past the first contact than the reference order), and a scorer
trained from scratch on GPU-logged samples beating the plain order on
held-out queries."""
import ctypes as C
import os

import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import learn as Lm
from c2a_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def small_knot(nq, seed):
    tris, idx = W.torus_knot(n_seg=128, n_ring=12)
    t9 = tris.reshape(-1, 9)
    return t9, [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)], W.knot_queries(nq, seed, 0.4)


def test_gpu_samples_match_cpu_harness():
    """The GPU logs the same samples (keys) as the kernel logic run on the
    CPU; features agree to BF16 rounding, labels to FP32 rounding."""
    t9, meshes, P = small_knot(12, 31)
    # the same traversal knobs on both sides: samples come from the replayed
    # records, whose set depends on the replay log and its extension
    knobs = dict(tol_d=1e-3, tol_t=1e-4, spec_depth=5, miss_depth=3, replay_log=8192,
                 ext_levels=0)
    params = G.default_params(sample_rate=1.0, sample_cap=300000, **knobs)
    b = G.Batch(meshes, G.make_queries(*P), params)
    b.run()
    X, y, k, n = b.samples(300000)
    assert n == len(k) > 500
    emu = C.CDLL(os.path.join(HERE, "emu", "libccd_emu.so"))
    emu.emu_set_sampling.argtypes = [C.c_double, C.c_int64]
    emu.emu_samples.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    emu.emu_samples.restype = C.c_int64
    emu.emu_set_sampling(1.0, 300000)
    try:
        from tests.test_kernel_logic_cpu import run_emu
        run_emu(emu, t9, t9, P, G.default_params(**knobs), 5)
        rows = np.zeros((300000, G.NF + 1), np.float32)
        keys = np.zeros(300000, np.int64)
        m = emu.emu_samples(rows.ctypes.data, keys.ctypes.data, 300000)
    finally:
        emu.emu_set_sampling(0.0, 0)
    o = np.argsort(keys[:m])
    keys, rows = keys[:m][o], rows[:m][o]
    assert np.array_equal(keys, k)
    scale = np.maximum(np.abs(rows[:, :G.NF]), 1e-30)
    assert (np.abs(X - rows[:, :G.NF]) <= scale * 2.0 ** -7).all()
    assert np.allclose(y, rows[:, G.NF], rtol=1e-6, atol=1e-6)


def test_gpu_sampling_changes_no_result():
    _, meshes, P = small_knot(64, 32)
    q = G.make_queries(*P)
    a = G.ccd_batch(meshes, q, G.default_params(tol_d=1e-3, tol_t=1e-4))
    b = G.Batch(meshes, q, G.default_params(tol_d=1e-3, tol_t=1e-4, sample_rate=0.5,
                                             sample_cap=1 << 20))
    b.run()
    r = b.fetch()
    for f in a.dtype.names:
        assert np.array_equal(a[f], r[f]), f


def overshoots(r):
    """Colliding results whose TOC pose is already in contact (toc > 0 and
    the exact closest distance there is 0): the CA loop stepped past the
    first contact.  The final step's distance is exact (the error schedule
    is 0 once the distance is below 0.5), so this needs no recomputation."""
    return int(((r["collisionfree"] == 0) & (r["toc"] > 0) & (r["distance"] == 0)).sum())


def check_learned(meshes, P, mp, scorer, tol_d, tol_t):
    """Learned order against plain order on the same queries: identical
    collision flags, and no more overshoots past the first contact than the
    reference order itself makes (plus 0.5 % of the colliding queries)."""
    q = G.make_queries(*P, mesh_pairs=mp)
    plain = G.ccd_batch(meshes, q, G.default_params(tol_d=tol_d, tol_t=tol_t, with_contacts=0))
    learned = G.ccd_batch(meshes, q, G.default_params(tol_d=tol_d, tol_t=tol_t, with_contacts=0,
                                                      order=G.ORDER_LEARNED, scorer=scorer))
    assert np.array_equal(plain["collisionfree"], learned["collisionfree"])
    ncol = int((plain["collisionfree"] == 0).sum())
    print("overshoots past first contact: plain %d, learned %d of %d colliding"
          % (overshoots(plain), overshoots(learned), ncol))
    assert overshoots(learned) <= overshoots(plain) + max(1, ncol // 200)
    return plain, learned


def test_gpu_learned_order_fewer_bv_tests_and_conservative():
    """The shipped scorer (trained on GPU logs of the knot and uneven
    batches) on the knot batch's first 1024 queries."""
    s = G.Scorer(path=Lm.DEFAULT_SCORER)
    tris, idx = W.torus_knot()
    t9 = tris.reshape(-1, 9)
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    P = [p[:1024] for p in W.knot_queries(4096, W.CONFIG_KNOT_BATCH["seed"], 0.3)]
    plain, learned = check_learned(meshes, P, None, s, 1e-3, 1e-4)
    c = Lm.compare(plain, learned)
    print("knot 1024:", c)
    assert c["bv_ratio"] > 1.15
    assert c["toc_within_1e-4"] > 0.8


def test_gpu_scorer_trained_from_scratch_gains_on_held_out_queries():
    """Samples logged by the plain GPU traversal on training queries, a
    scorer trained on them by the matrix-core kernels, and fewer BV tests
    than the plain order on queries it never saw."""
    t9, meshes, P = small_knot(1024, 33)
    q = G.make_queries(*P)
    X, y, k, n = Lm.collect_samples(meshes, q, 1e-3, 1e-4, 0.2, 4_000_000)
    assert 0 < n <= 4_000_000
    train = Lm.split_mask(k, 0.7, 3)
    s = Lm.train_scorer(X[train], y[train], epochs=6, seed=2)
    sc = s.infer(X[~train])
    acc_l, npair = Lm.sibling_accuracy(k[~train], y[~train], sc)
    acc_p, _ = Lm.sibling_accuracy(k[~train], y[~train], X[~train, 18])
    print("held-out sibling accuracy: learned %.4f plain %.4f over %d pairs" % (acc_l, acc_p, npair))
    assert acc_l > acc_p
    held = ~Lm.query_split(1024, 0.7, 3)
    Ph = [p[held] for p in P]
    plain, learned = check_learned(meshes, Ph, None, s, 1e-3, 1e-4)
    c = Lm.compare(plain, learned)
    print("held-out queries:", c)
    assert c["bv_ratio"] > 1.05
