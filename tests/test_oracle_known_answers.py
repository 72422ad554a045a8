"""Known answers for the oracle's order-dependent parts, derived by hand.

* The CA loop (C2A.cpp:2029-2143) driven by a scripted step: every branch of
  the loop -- convergence, the first-step exit (toc = first mint), the
  collision-free exits, the tol_t break, the iteration cap and the
  toc >= 1 - tol_t rule -- with toc, numCA, flags and the lamda /
  UpboundTOC / numCA each step call sees written out by hand.
* The hierarchy build (C2A_Build.cpp:344-488) on a mesh whose covariance is
  diagonal at every node, so the eigenvectors, the axis order, the split
  axis, every split (the x <= c tie, the swap order of the partition, the
  half-split fallback) and the node numbering are known without running the
  build.  The product's builder (c2a_mesh_export) must give the same tree.
* One CA step (C2A.cpp:1778-1931 -> 1113-1354) over two 2-triangle models in
  pure translation, whose RSS distances, visiting order, prunes, improving
  leaf, mint, counters and last_tri are computed by hand.

Every expected value below is a literal or an expression in the test's own
arithmetic; none is computed by the oracle.
"""
import numpy as np
import pytest

import c2a_amd as G
from oracle import oracle as O

TOL_D, TOL_T = 1e-3, 1e-4


# ---------------------------------------------------------------------------
# CA loop, scripted step
# ---------------------------------------------------------------------------
def loop(dist, mint, max_iter=150):
    return O.ca_loop_scripted(dist, mint, TOL_D, TOL_T, max_iter)


def test_ca_loop_converges():
    res, calls, lam, up, nca = loop([0.5, 0.1, 0.01, 5e-4], [0.2, 0.3, 0.1, 0.05])
    assert calls == 4
    assert list(lam) == [0.0, 0.2, 0.2 + 0.3, (0.2 + 0.3) + 0.1]
    assert list(up) == [1.0, 1.0 - 0.2, 1.0 - (0.2 + 0.3), 1.0 - ((0.2 + 0.3) + 0.1)]
    assert list(nca) == [0, 2, 3, 4]          # numCA is incremented before the step
    assert res.toc == (0.2 + 0.3) + 0.1 and res.numCA == 4
    assert res.collisionfree == 0 and res.flags == 0


def test_ca_loop_first_step_within_tolerance_returns_first_mint():
    res, calls, *_ = loop([5e-4], [0.03])     # C2A.cpp:2072: lastLamda = mint
    assert calls == 1 and res.toc == 0.03 and res.numCA == 1 and res.collisionfree == 0


def test_ca_loop_first_mint_one_is_collision_free():
    res, calls, *_ = loop([0.7], [1.0])
    assert calls == 1 and res.collisionfree == 1 and res.toc == 0.0 and res.numCA == 1


def test_ca_loop_lamda_past_one_is_collision_free():
    res, calls, lam, *_ = loop([0.5, 0.5], [0.6, 0.5])
    assert calls == 2 and list(lam) == [0.0, 0.6]
    assert res.collisionfree == 1 and res.toc == 0.0 and res.numCA == 2


def test_ca_loop_small_mint_breaks_with_last_lamda():
    res, calls, *_ = loop([0.5, 0.3], [0.25, 5e-5])
    assert calls == 2 and res.toc == 0.25 and res.collisionfree == 0 and res.numCA == 2


def test_ca_loop_mint_equal_to_tol_t_still_advances():
    res, calls, lam, *_ = loop([0.5, 5e-4], [1e-4, 0.5])
    assert calls == 2 and list(lam) == [0.0, 1e-4] and res.toc == 1e-4


def test_ca_loop_iteration_cap():
    res, calls, lam, *_ = loop([0.5], [1e-3], max_iter=150)
    acc = 0.0
    for _ in range(150):
        acc += 1e-3
    assert calls == 151 and res.numCA == 151
    assert res.flags & O.FLAG_ITER_CAP and res.collisionfree == 0
    assert res.toc == acc and lam[-1] == acc


def test_ca_loop_toc_near_one_reports_zero():
    res, calls, *_ = loop([0.5, 5e-4], [0.99995, 0.3])
    assert calls == 2 and res.toc == 0.0 and res.collisionfree == 0
    res, calls, *_ = loop([5e-4], [1.0])      # first-step exit with mint 1
    assert calls == 1 and res.toc == 0.0 and res.collisionfree == 0


# ---------------------------------------------------------------------------
# hierarchy build on a mesh with diagonal covariance at every node
# ---------------------------------------------------------------------------
H = 0.125


def slab_tri(x):
    """A triangle in the plane X = x with vertices (x, H, H), (x, -H, H),
    (x, 0, -2H): per triangle sum y = sum z = sum yz = 0, and with a common
    x also sum xy = sum xz = 0, so the covariance of any set of these
    triangles is exactly diag(var_x, 2H^2 k, 6H^2 k)."""
    return np.array([[x, H, H], [x, -H, H], [x, 0.0, -2 * H]], dtype=np.float64).reshape(9)


# Triangles by id (AddTri order) and their x.  Hand-derived build:
#  root {t0..t4}: x = 2,0,5,2,1, var_x = 42 > var_z > var_y: split axis x,
#    c = mean x = 2; partition (swap order): t0 L, t1 L, t2 R, t3 L (x == c,
#    the tie goes left), t4 L -> [t0 t1 t3 t4 | t2]; children nodes 1, 2.
#  node 1 {t0 t1 t3 t4}: c = 1.25 -> t1, t4 left (swaps) -> [t1 t4 | t3 t0];
#    children 3, 4.
#  node 3 {t1 t4}: c = 0.5 -> [t1 | t4]; children 5, 6 (leaves).
#  node 4 {t3 t0}: both x = 2, var_x = 0 < var_y < var_z: split axis z; both
#    centroids have z = 0 = c, so every triangle lands left and the half
#    split takes over -> [t3 | t0]; children 7, 8 (leaves).
#  node 2 = t2 (leaf).
XS = [2.0, 0.0, 5.0, 2.0, 1.0]
BUILD_ORDER = [1, 4, 3, 0, 2]                     # triangle ids in build order
FIRST_CHILD = [1, 3, -5, 5, 7, -1, -2, -3, -4]    # leaves: -(build index + 1)
DEPTH = [0, 1, 1, 2, 2, 3, 3, 3, 3]
# model-frame orientation of the two split nodes: columns = eigenvectors by
# decreasing eigenvalue, third column = first x second (C2A_Build.cpp:426-430)
R_ROOT = np.array([[1.0, 0, 0], [0, 0, -1.0], [0, 1.0, 0]])     # x, z, x cross z
R_NODE4 = np.array([[0, 0, -1.0], [0, 1.0, 0], [1.0, 0, 0]])    # z, y, z cross y


def slab_mesh():
    return np.stack([slab_tri(x) for x in XS])


def test_build_known_tree_oracle():
    m = O.OModel(slab_mesh(), ids=np.arange(5, dtype=np.int32))
    bv = m.bv_arrays()
    assert list(bv["first_child"]) == FIRST_CHILD
    assert list(bv["depth"]) == DEPTH
    assert list(m.tris()["id"]) == BUILD_ORDER
    assert np.array_equal(np.array(bv[0]["R_loc"]), R_ROOT)
    assert np.array_equal(np.array(bv[4]["R_loc"]), R_NODE4)
    assert m.last_tri == 0


def test_build_known_tree_product():
    m = G.Mesh(slab_mesh(), ids=np.arange(5, dtype=np.int32))
    nodes, tris = m.export()
    assert [int(v) for v in nodes[:, 26]] == FIRST_CHILD
    assert [int(v) for v in nodes[:, 27]] == DEPTH
    assert list(m.tri_ids()) == BUILD_ORDER
    assert np.array_equal(tris, slab_mesh()[BUILD_ORDER])
    assert np.array_equal(nodes[0, 16:25].reshape(3, 3), R_ROOT)
    assert np.array_equal(nodes[4, 16:25].reshape(3, 3), R_NODE4)


# ---------------------------------------------------------------------------
# one CA step over two 2-triangle models, by hand
# ---------------------------------------------------------------------------
# Model 1: slab triangles at x = 0 (A0) and x = 4 (A1), at rest at the
# origin.  Model 2: slab triangles at x = 0 (B0) and x = 2 (B1), starting at
# X = 10 and moving by (-16, 0, 0) over the unit interval, no rotation.
# Build: each model splits along x at its mean -> build order A0, A1 / B0, B1.
# RSS volumes: a leaf is its triangle's rectangle in the plane X = x
# (radius 0); a root spans x (length 4 or 2) with radius H about y = 0.
# Sizes: model 1's root is larger, so its children are visited first
# (C2A.cpp:1195): pairs (A0, Broot) at d = 10 - H and (A1, Broot) at
# d = 6 - H.  Motion bounds: object 1 is at rest (path 1e-30), object 2
# approaches at 16, so mint = d / (1e-30 + 16) = d / 16.
# With numCA > 5 the error schedule is zero: descend iff mint < UpboundTOC and
# d < distance.
def two_tri_models():
    m1 = np.stack([slab_tri(0.0), slab_tri(4.0)])
    m2 = np.stack([slab_tri(0.0), slab_tri(2.0)])
    return m1, m2


def step(seeds):
    a, b = two_tri_models()
    M1, M2 = O.OModel(a), O.OModel(b)
    I, z = np.eye(3), np.zeros(3)
    mo1 = O.OMotion(I, z, I, z)
    mo2 = O.OMotion(I, np.array([10.0, 0, 0]), I, np.array([-6.0, 0, 0]))
    return O.toc_step(M1, M2, mo1, mo2, I, z, I, np.array([10.0, 0, 0]), seeds, 6, 1.0, 1.0), M1, M2


def test_two_triangle_step_from_far_seed():
    # seed (A0, B0): distance 10.  Visit (A1, Broot) first (6 - H < 10 - H):
    # descend (d = 5.875 < 10); it splits model 2: (A1, B0) at 6, (A1, B1) at
    # 8 -> (A1, B0) first (not d2 < d1), descend, leaf 6 <= 10 improves:
    # distance 6, mint 6 / 16, last (A1, B0) = (1, 0).  (A1, B1): 8 >= 6,
    # pruned with mint 8 / 16.  (A0, Broot): 9.875 >= 6, pruned, mint
    # 9.875 / 16.  BV tests: 2 per expanded pair (the root and (A1, Broot);
    # a leaf pair is tested, not expanded) = 4; one triangle test.
    res, M1, M2 = step((0, 0))
    assert res.distance == 6.0
    assert res.mint == 6.0 / 16.0
    # res->R, res->T: model 2's pose in model 1's frame (C2A.cpp:1793-1796)
    assert np.array_equal(np.array(res.R), np.eye(3)) and list(res.T) == [10.0, 0.0, 0.0]
    assert (res.num_bv_tests, res.num_tri_tests) == (4, 1)
    assert (M1.last_tri, M2.last_tri) == (1, 0)
    assert (res.abs_err, res.rel_err) == (0.0, 0.0)


def test_two_triangle_step_from_near_seed():
    # seed (A1, B0): distance 6.  (A1, Broot) at 5.875 < 6: descend; (A1, B0)
    # at 6 is not < 6: pruned (mint 6 / 16); (A1, B1) pruned (8 / 16);
    # (A0, Broot) pruned.  No leaf is tested; mint = 6 / 16 from a pruned
    # pair; last_tri unchanged.
    res, M1, M2 = step((1, 0))
    assert res.distance == 6.0 and res.mint == 6.0 / 16.0
    assert (res.num_bv_tests, res.num_tri_tests) == (4, 0)
    assert (M1.last_tri, M2.last_tri) == (0, 0)


def test_two_triangle_step_upbound_prunes():
    # UpboundTOC 0.3: (A1, Broot) has mint 5.875 / 16 = 0.367 >= 0.3, so it is
    # pruned at once; (A0, Broot) too.  mint = 5.875 / 16.
    a, b = two_tri_models()
    M1, M2 = O.OModel(a), O.OModel(b)
    I, z = np.eye(3), np.zeros(3)
    mo1 = O.OMotion(I, z, I, z)
    mo2 = O.OMotion(I, np.array([10.0, 0, 0]), I, np.array([-6.0, 0, 0]))
    res = O.toc_step(M1, M2, mo1, mo2, I, z, I, np.array([10.0, 0, 0]), (0, 0), 6, 1.0, 0.3)
    assert res.distance == 10.0 and res.mint == (6.0 - H) / 16.0
    assert (res.num_bv_tests, res.num_tri_tests) == (2, 0)


@pytest.mark.parametrize("strict", [0, 1])
def test_two_triangle_step_strict_leaf_rule(strict):
    # From the far seed the only tested leaf improves, so strict mode gives
    # the same step; the world axis equals the body axis (no rotation).
    with O.strict_mode(bool(strict)):
        res, M1, M2 = step((0, 0))
    assert (res.distance, res.mint, res.num_bv_tests) == (6.0, 6.0 / 16.0, 4)
