"""c2a_ccd_batch_noalloc, the allocation-free entry point, through the C ABI.

The caller owns every device buffer (torch tensors here, as plain device
memory); the library's own device memory is its fixed per-device scratch.
Results are compared with c2a_ccd_batch on the same queries for exact
equality, on two calls in a row (the second with a smaller input, so stale
scratch contents from the first call would show), in the reference mode and
in strict mode; out-of-range mesh ids and seeds come back flagged.
"""
import numpy as np
import pytest
import torch

import c2a_amd as G
from c2a_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ["toc", "distance", "mint", "R_toc", "T_toc", "p1", "p2", "bv_tests", "tri_tests",
          "numCA", "collisionfree", "num_contact", "last_tri", "flags"]


@pytest.fixture(scope="module")
def knot_meshes():
    tris, idx = W.torus_knot(n_seg=128, n_ring=12)
    t9 = tris.reshape(-1, 9)
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    # meshes reach the device with their first allocating launch
    G.ccd_batch(meshes, G.make_queries(*W.knot_queries(1, 60)), G.default_params())
    return t9, meshes


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run_noalloc(meshes, P, params, pairs=None, seeds=None):
    nq = len(P[0])
    mot = dev(G.motion_records(*P))
    mp = dev(np.tile(np.array([[0, 1]], np.int32), (nq, 1)) if pairs is None else pairs)
    sd = None if seeds is None else dev(seeds)
    out = torch.zeros(nq * G.RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    G.ccd_batch_noalloc(meshes, mot.data_ptr(), mp.data_ptr(),
                        None if sd is None else sd.data_ptr(), nq, params, out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().view(G.RESULT_DTYPE).copy()


def assert_equal(ref, got, label):
    for f in FIELDS:
        assert np.array_equal(ref[f], got[f]), (label, f, np.nonzero(ref[f] != got[f])[0][:5])


@pytest.mark.parametrize("strict", [0, 1])
def test_noalloc_equals_batch_two_calls(knot_meshes, strict):
    t9, meshes = knot_meshes
    params = G.default_params(tol_d=1e-3, tol_t=1e-4, strict=strict)
    for nq, seed in ((192, 61), (40, 62)):
        print("noalloc strict=%d nq=%d" % (strict, nq), flush=True)
        P = W.knot_queries(nq, seed, collide_frac=0.4)
        seeds = np.stack([np.arange(nq) % len(t9), (5 * np.arange(nq)) % len(t9)], 1).astype(np.int32)
        ref = G.ccd_batch(meshes, G.make_queries(*P, seeds=seeds), params)
        got = run_noalloc(meshes, P, params, seeds=seeds)
        assert_equal(ref, got, "nq=%d" % nq)
        assert (ref["collisionfree"] == 0).any() and (ref["collisionfree"] == 1).any()


def test_noalloc_no_replay_and_bad_inputs(knot_meshes):
    t9, meshes = knot_meshes
    nq = 24
    P = W.knot_queries(nq, 63, collide_frac=0.4)
    params = G.default_params(tol_d=1e-3, tol_t=1e-4, replay_log=0)
    ref = G.ccd_batch(meshes, G.make_queries(*P), params)
    pairs = np.tile(np.array([[0, 1]], np.int32), (nq, 1))
    seeds = np.zeros((nq, 2), np.int32)
    pairs[3] = [0, 2]            # mesh id out of range
    seeds[7] = [0, len(t9)]      # seed triangle out of range
    got = run_noalloc(meshes, P, params, pairs=pairs, seeds=seeds)
    good = np.ones(nq, bool)
    good[[3, 7]] = False
    assert_equal(ref[good], got[good], "replay_log=0")
    assert (got["flags"][~good] == G.FLAG_BAD_INPUT).all()
    assert (got["numCA"][~good] == 0).all()


def test_noalloc_rejects_sizes_beyond_scratch(knot_meshes):
    _, meshes = knot_meshes
    with pytest.raises(G.C2AError) as e:
        G.ccd_batch_noalloc(meshes, 1, 1, None, 4, G.default_params(replay_log=1 << 20), 1)
    assert e.value.code == G.ERR_BAD_ARG and "scratch" in str(e.value)
    with pytest.raises(G.C2AError) as e:
        G.ccd_batch_noalloc(meshes * 17, 1, 1, None, 4, None, 1)
    assert e.value.code == G.ERR_BAD_ARG and "at most 32" in str(e.value)
