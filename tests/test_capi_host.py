"""Host side of the C ABI (no GPU needed): exports, the hierarchy builder,
motion preparation and error statuses, each against the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O


def test_library_exports_every_declared_symbol():
    header = open(G.HEADER_PATH).read()
    declared = set(re.findall(r"\b(c2a_[a-z_0-9]+)\s*\(", header))
    lib = C.CDLL(G.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(G.EXPORTS) == declared


def _compare_tree(tris, idx=None):
    m = G.Mesh(tris.reshape(-1, 9), vidx=idx)
    nodes, t9 = m.export()
    om = O.OModel(tris.reshape(-1, 9), idx=idx)
    bv = om.bv_arrays()
    assert m.nbv == om.nbv and m.max_depth == bv["depth"].max()
    assert np.array_equal(t9, om.tri_xyz().reshape(-1, 9))
    assert np.array_equal(nodes[:, 0:9], bv["R"].reshape(-1, 9))
    assert np.array_equal(nodes[:, 9:12], bv["Tr"])
    assert np.array_equal(nodes[:, 12:14], bv["l"])
    assert np.array_equal(nodes[:, 14], bv["r"])
    assert np.array_equal(nodes[:, 15], bv["angR"])
    assert np.array_equal(nodes[:, 16:25], bv["R_loc"].reshape(-1, 9))
    assert np.array_equal(nodes[:, 25], bv["size"])
    assert np.array_equal(nodes[:, 26].astype(np.int64), bv["first_child"])
    assert np.array_equal(nodes[:, 27].astype(np.int64), bv["depth"])
    ids = m.tri_ids()
    assert sorted(ids) == list(range(len(tris)))


@pytest.mark.parametrize("shape", ["knot", "torus", "ico3", "uv_small"])
def test_host_hierarchy_is_bit_exact(shape):
    if shape == "knot":
        tris, idx = W.torus_knot()
    elif shape == "torus":
        tris, idx = W.torus()
    elif shape == "ico3":
        tris, idx = W.icosphere(3, 0.1, 5)
    else:
        tris, idx = W.uv_sphere(64, 32, noise_amp=0.1, seed=2)
    _compare_tree(tris, idx)


def test_motion_preparation_and_poses_are_bit_exact():
    rng = np.random.default_rng(11)
    P = W.knot_queries(64, 99)
    for i in range(64):
        m = G.motion_prepare(P[0][i], P[1][i])
        om = O.OMotion(P[0][i][:9].reshape(3, 3), P[0][i][9:], P[1][i][:9].reshape(3, 3), P[1][i][9:])
        assert np.array_equal(m[12:16], np.array(om.m.q0[:]))
        assert np.array_equal(m[16:19], om.cv) and np.array_equal(m[19:22], om.axis)
        assert m[22] == om.angVel
        for t in np.concatenate([[0.0, 1.0, 1.5], rng.uniform(0, 1, 5)]):
            pose, q = G.motion_pose(m, t)
            R, T, _ = om.integrate(t)
            assert np.array_equal(pose[:9], R.reshape(-1)) and np.array_equal(pose[9:], T)


def test_error_statuses():
    L = G.lib()
    h = C.c_void_p()
    rc = L.c2a_mesh_create(None, None, None, 0, C.byref(h))
    assert rc == G.ERR_EMPTY_MODEL
    rc = L.c2a_mesh_create(None, None, None, 5, C.byref(h))
    assert rc == G.ERR_BAD_ARG
    assert L.c2a_mesh_info(None, None, None, None) == G.ERR_BAD_ARG
    assert L.c2a_status_string(G.ERR_HIP) == b"HIP runtime error"
    with pytest.raises(G.C2AError):
        G.Mesh(np.zeros((0, 9)))
    p = G.default_params()
    assert (p.tol_d, p.tol_t, p.max_iter, p.with_contacts) == (1e-4, 1e-4, 150, 1)
    assert (p.contact_scale, p.contact_offset) == (2.0, 0.001)


def test_batch_rejects_bad_queries():
    tris, idx = W.icosphere(1)
    m = G.Mesh(tris.reshape(-1, 9))
    P = W.knot_queries(2, 1)
    q = G.make_queries(*P, mesh_pairs=[[0, 3], [0, 0]])
    with pytest.raises(G.C2AError) as e:
        G.ccd_batch([m], q)
    assert e.value.code in (G.ERR_BAD_ARG, G.ERR_NO_DEVICE)
