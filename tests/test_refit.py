"""Refit (c2a_mesh_refit): the tree, the triangle order and the node numbers
stay, and every node's RSS volume is fitted again to moved triangles with
the build's fit.  The reference has no refit (it rebuilds), so there is no
result to be exact against; the checks are what a refit must satisfy:

- every triangle of a node lies inside the node's volume after the move
  (a wrong triangle range, a stale frame or a mixed-up parent transform
  breaks this);
- the refit depends only on the triangles: refitting to the moved mesh and
  back gives the first refit bit for bit;
- the GPU refit equals the host refit bit for bit;
- exact distances over a refitted mesh equal the oracle's over the moved
  triangles (the tree differs, the minimum does not)."""
import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W


def deform(tris, amp=0.05, seed=0):
    """A smooth deformation of the vertices (shared vertices stay shared)."""
    p = tris.reshape(-1, 3).copy()
    rng = np.random.default_rng(seed)
    k = rng.normal(size=(3, 3)) * 3.0
    ph = rng.uniform(0, 2 * np.pi, 3)
    for a in range(3):
        p[:, a] += amp * np.sin(p @ k[a] + ph[a])
    return p.reshape(-1, 9)


def node_ranges(fc):
    n = len(fc)
    first = np.zeros(n, np.int64)
    num = np.zeros(n, np.int64)
    for v in range(n - 1, -1, -1):
        c = int(fc[v])
        if c < 0:
            first[v], num[v] = -c - 1, 1
        else:
            first[v], num[v] = first[c], num[c] + num[c + 1]
    return first, num


def model_frames(nodes):
    """Model-frame corner of every node from the parent-relative frames."""
    fc = nodes[:, 26].astype(np.int64)
    n = len(nodes)
    T = np.zeros((n, 3))
    T[0] = nodes[0, 9:12]
    for v in range(n):
        c = int(fc[v])
        if c >= 0:
            Rp = nodes[v, 16:25].reshape(3, 3)
            for ch in (c, c + 1):
                T[ch] = Rp @ nodes[ch, 9:12] + T[v]
    return T


def check_contains(nodes, tris, tol=1e-9):
    first, num = node_ranges(nodes[:, 26].astype(np.int64))
    T = model_frames(nodes)
    worst = 0.0
    for v in range(len(nodes)):
        R = nodes[v, 16:25].reshape(3, 3)
        l0, l1, r = nodes[v, 12], nodes[v, 13], nodes[v, 14]
        pts = tris[first[v]:first[v] + num[v]].reshape(-1, 3)
        q = (pts - T[v]) @ R   # rectangle frame
        dx = np.maximum(0, np.maximum(-q[:, 0], q[:, 0] - l0))
        dy = np.maximum(0, np.maximum(-q[:, 1], q[:, 1] - l1))
        d = np.sqrt(dx * dx + dy * dy + q[:, 2] ** 2) - r
        worst = max(worst, float(d.max()))
    assert worst <= tol, worst
    return worst


@pytest.fixture(scope="module")
def sphere():
    t, idx = W.icosphere(3, 0.1, 5)   # 1280 triangles
    return t.reshape(-1, 9), idx


def test_build_volumes_contain_their_triangles(sphere):
    """The containment check on a plain build (it must hold there)."""
    t9, idx = sphere
    m = G.Mesh(t9, vidx=idx)
    nodes, tris = m.export()
    check_contains(nodes, tris)


def test_refit_contains_moved_triangles_and_depends_only_on_them(sphere):
    t9, idx = sphere
    m = G.Mesh(t9, vidx=idx)
    n0, _ = m.export()
    moved = deform(t9, 0.08, 1)
    m.refit(moved, device=False)
    n1, tr1 = m.export()
    # the tree is kept
    assert np.array_equal(n0[:, 26], n1[:, 26]) and np.array_equal(n0[:, 27], n1[:, 27])
    # the triangles are the moved ones, in build order
    first, _ = node_ranges(n1[:, 26].astype(np.int64))
    assert np.array_equal(np.sort(tr1, axis=0), np.sort(moved, axis=0))
    check_contains(n1, tr1)
    # the volumes changed: the tree no longer bounds the old positions
    assert not np.array_equal(n0[:, 12:16], n1[:, 12:16])
    # back to the original triangles and forward again: the same bits
    m.refit(t9, device=False)
    nb, _ = m.export()
    check_contains(nb, m.export()[1])
    m.refit(moved, device=False)
    n2, tr2 = m.export()
    assert np.array_equal(n1, n2) and np.array_equal(tr1, tr2)


def test_refit_at_unchanged_triangles_is_close_to_the_build(sphere):
    """Same triangles: the refit fits each node over its triangles in their
    final order, the build over the order before that node's split, so the
    two agree up to the order of the sums (and of the corner pass)."""
    t9, idx = sphere
    m = G.Mesh(t9, vidx=idx)
    n0, t0 = m.export()
    m.refit(t9, device=False)
    n1, t1 = m.export()
    assert np.array_equal(t0, t1)
    # orientations and radii agree to rounding; the rectangles can differ by
    # where the corner pass grew them (both contain every triangle)
    assert np.allclose(n0[:, 0:9], n1[:, 0:9], rtol=0, atol=1e-11)
    assert np.allclose(n0[:, 16:25], n1[:, 16:25], rtol=0, atol=1e-11)
    assert np.allclose(n0[:, 14], n1[:, 14], rtol=0, atol=1e-12)
    assert np.array_equal(n0[:, 15], n1[:, 15])   # angular radius: a maximum
    check_contains(n1, t1)
    same = (n0 == n1).all(1).mean()
    assert same > 0.4, same


def test_refit_rejects_wrong_sizes(sphere):
    t9, idx = sphere
    m = G.Mesh(t9, vidx=idx)
    with pytest.raises(ValueError):
        m.refit(t9[:-1], device=False)


# ---- GPU -------------------------------------------------------------------

def _gpu_meshes():
    t, idx = W.torus_knot()
    yield "knot 8192", t.reshape(-1, 9), idx
    t, idx = W.icosphere(6, 0.05, 5)
    yield "icosphere 81920", t.reshape(-1, 9), idx
    s, idx = W.uv_sphere(512, 256, 0.0, 3)
    yield "uv sphere 262144", s.reshape(-1, 9), idx


@pytest.mark.gpu
def test_device_refit_equals_host_refit():
    for name, t9, idx in _gpu_meshes():
        print("refit: %s" % name, flush=True)
        moved = deform(t9, 0.05, 2)
        h = G.Mesh(t9, vidx=idx)
        d = G.Mesh(t9, vidx=idx, device=True)
        h.refit(moved, device=False)
        d.refit(moved, device=True)
        hn, ht = h.export()
        dn, dt = d.export()
        bad = np.nonzero((hn != dn).any(1))[0]
        assert len(bad) == 0, (name, len(bad), bad[:5])
        assert np.array_equal(ht, dt), name


@pytest.mark.gpu
def test_distances_over_a_refitted_mesh_are_exact():
    """Exact static distances (rel_err = abs_err = 0) over the refitted knot
    equal the oracle's over the moved triangles with its own, rebuilt tree."""
    from oracle import oracle as O
    t, idx = W.torus_knot()
    t9 = t.reshape(-1, 9)
    moved = deform(t9, 0.03, 3)
    m = G.Mesh(t9, vidx=idx, device=True)
    m.refit(moved, device=True)
    P = W.knot_queries(256, 99, 0.0)
    poses = np.concatenate([P[0], P[2]], 1)
    got = G.distance([m, m], poses, np.tile([0, 1], (len(poses), 1)))
    om = O.OModel(moved)
    ref = np.array([r.distance for r in O.distance_batch(om, om, poses, nthreads=8)])
    assert (ref > 0).any()
    assert np.array_equal(got["distance"], ref)
