"""Strict (conservative) mode, on the CPU: the oracle's strict switches and the
kernel logic (tests/emu) against them.

The reference's CA step is not conservative in two places:
  * both motion bounds cross the body-frame rotation axis with world-frame
    directions (InterpMotion.cpp:784-791, 838-845 with C2A.cpp:1154-1164,
    1240-1242), which bounds the rotation only when the start rotation is I;
  * only improving leaf pairs (dTri <= distance, C2A.cpp:1150) bound mint.
Strict mode crosses the world-frame axis R0 * axis, lets every tested leaf
pair bound mint (a touching one by 0), and must then never report a TOC at
or after the first contact.  Each check below fails on the reference
semantics or on a plausible mistake in the switches.
"""
import ctypes as C

import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O
from tests import brute as B
from tests.test_kernel_logic_cpu import emu, run_emu, same  # noqa: F401


def rng(seed):
    return np.random.default_rng(seed)


def _disp_ratio(m, angR, N, r):
    """Largest sampled displacement along n over bound * t, for points within
    angR of the model origin moved by motion m from several start times."""
    n = N / np.linalg.norm(N)
    pts = r.normal(size=(96, 3))
    pts = pts / np.linalg.norm(pts, axis=1)[:, None] * r.uniform(0.5, 1.0, (96, 1)) * angR
    bound = m.bound(angR, N)
    worst = 0.0
    for lam in (0.0, 0.3, 0.7):
        Rl, Tl, _ = m.integrate(lam)
        for dt in (0.05, 0.2, 0.3):
            Rt, Tt, _ = m.integrate(lam + dt)
            disp = (pts @ Rt.T + Tt - (pts @ Rl.T + Tl)) @ n
            worst = max(worst, disp.max() / (bound * dt))
    return worst


def test_world_axis_is_R0_axis():
    r = rng(21)
    for _ in range(20):
        R0, R1 = B.random_rotation(r), B.random_rotation(r)
        m = O.OMotion(R0, np.zeros(3), R1, np.zeros(3))
        assert np.array_equal(np.array(m.m.axw[:]), R0 @ m.axis) or \
            np.allclose(np.array(m.m.axw[:]), R0 @ m.axis, atol=1e-15)
        # the world axis is the fixed axis of the relative rotation R(t) R0^T
        Rt, _, _ = m.integrate(0.6)
        assert np.allclose((Rt @ R0.T) @ np.array(m.m.axw[:]), m.m.axw[:], atol=1e-12)


def test_strict_bounds_hold_for_any_start_rotation():
    """Sampled displacement along the bound's direction never exceeds
    bound * dt in strict mode, start rotation random (R0 != I); the
    reference's body-frame bound is exceeded on the same samples."""
    r = rng(22)
    worst_ref = worst_strict = 0.0
    for _ in range(120):
        R0, R1 = B.random_rotation(r), B.random_rotation(r)
        m = O.OMotion(R0, np.zeros(3), R1, r.normal(size=3) * 0.05)
        N = r.normal(size=3)
        angR = r.uniform(0.5, 2.0)
        with O.strict_mode(True):
            s = _disp_ratio(m, angR, N, np.random.default_rng(7))
        with O.strict_mode(False):
            f = _disp_ratio(m, angR, N, np.random.default_rng(7))
        worst_strict, worst_ref = max(worst_strict, s), max(worst_ref, f)
    assert worst_strict <= 1.0 + 1e-9
    assert worst_ref > 1.05     # the reference quirk is real on these motions


def test_strict_leaf_bound_uses_world_axis():
    """The leaf bound (computeTOC) takes the same axis switch."""
    r = rng(23)
    R0 = B.rodrigues([1, 0, 0], 1.2)
    R1 = R0 @ B.rodrigues([0, 0, 1], 0.8)        # body-frame rotation about z
    m = O.OMotion(R0, np.zeros(3), R1, np.zeros(3))
    assert np.allclose(m.axis, [0, 0, 1], atol=1e-12)
    S = R0 @ np.array([0.0, 0.0, 1.0])           # world direction of the body z axis
    with O.strict_mode(True):
        assert m.bound_leaf(1.0, S) == pytest.approx(1e-30, abs=1e-12)   # turning about S
    with O.strict_mode(False):
        assert m.bound_leaf(1.0, S) > 0.5        # body z crossed with a world vector


def test_strict_mint_below_reference_when_axes_agree(spheres_pair):
    """With R0 = I for both objects the two axis choices coincide, so strict
    mode differs only in which leaves bound mint: same decisions, same
    distance and counters, and mint never larger than the reference's."""
    t1, t2 = spheres_pair
    m1, m2 = O.OModel(t1), O.OModel(t2)
    r = rng(24)
    I = np.eye(3)
    lower = 0
    for k in range(24):
        mo1 = O.OMotion(I, np.zeros(3), B.random_rotation(r), r.normal(size=3) * 0.3)
        mo2 = O.OMotion(I, np.array([2.3, 0, 0]) + r.normal(size=3) * 0.1, B.random_rotation(r),
                        np.array([0.5, 0, 0]))
        out = []
        for strict in (False, True):
            with O.strict_mode(strict):
                R1, T1, _ = mo1.integrate(0.0)
                R2, T2, _ = mo2.integrate(0.0)
                res = O.toc_step(m1, m2, mo1, mo2, R1, T1, R2, T2, (0, 0), 9, 0.0, 1.0)
                out.append((res.distance, res.num_bv_tests, res.num_tri_tests, res.mint))
        (dr, br, tr, mr), (ds, bs, ts, ms) = out
        assert (dr, br, tr) == (ds, bs, ts), k
        assert ms <= mr, k
        lower += ms < mr
    assert lower > 0     # the extra leaves matter on some of these steps


@pytest.fixture(scope="module")
def spheres_pair():
    t1, _ = W.icosphere(2, 0.1, 31)
    t2, _ = W.icosphere(2, 0.1, 32)
    return t1.reshape(-1, 9), t2.reshape(-1, 9)


def _exact_distance(m1, m2, P, i, t):
    mo1 = O.OMotion(P[0][i, :9].reshape(3, 3), P[0][i, 9:], P[1][i, :9].reshape(3, 3), P[1][i, 9:])
    mo2 = O.OMotion(P[2][i, :9].reshape(3, 3), P[2][i, 9:], P[3][i, :9].reshape(3, 3), P[3][i, 9:])
    Ra, Ta, _ = mo1.integrate(t)
    Rb, Tb, _ = mo2.integrate(t)
    return O.distance(m1, m2, Ra, Ta, Rb, Tb, 0.0, 0.0, seeds=(0, 0)).distance


def test_strict_never_reaches_contact_knot():
    """A 768-query slice of the seeded knot stream knot_queries(4096, 20261,
    0.3): in strict mode no colliding query's TOC pose, nor any of 8 poses
    sampled before it, is in contact (exact distance > 0); the reference
    semantics reach contact on this slice (query 608 of the stream)."""
    tris, _ = W.torus_knot()
    t9 = tris.reshape(-1, 9)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    P = W.knot_queries(4096, W.CONFIG_KNOT_BATCH["seed"], collide_frac=0.3)
    P = [p[256:1024] for p in P]
    nq = len(P[0])
    seeds = np.zeros((nq, 2), np.int32)
    with O.strict_mode(False):
        ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=seeds, nthreads=8)
    with O.strict_mode(True):
        res = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=seeds, nthreads=8)
    hit_ref = [i for i in np.nonzero((ref["collisionfree"] == 0) & (ref["toc"] > 0))[0]
               if _exact_distance(m1, m2, P, i, ref["toc"][i]) == 0.0]
    assert hit_ref, "the reference semantics reach contact on this slice"
    coll = np.nonzero((res["collisionfree"] == 0) & (res["toc"] > 0))[0]
    assert len(coll) > 300
    for i in coll:
        toc = res["toc"][i]
        for t in np.linspace(0.0, toc, 9):
            assert _exact_distance(m1, m2, P, i, t) > 0.0, (i, t, toc)


def test_strict_toc_before_bisected_contact(spheres_pair):
    """Head-on and glancing approaches of two noisy icospheres with random
    start rotations: the strict TOC is at or before the first contact found
    by bisection on the exact (brute-force-pinned) distance."""
    t1, t2 = spheres_pair
    m1, m2 = O.OModel(t1), O.OModel(t2)
    r = rng(25)
    n_hit = 0
    with O.strict_mode(True):
        for k in range(10):
            R00, R01, R10, R11 = (B.random_rotation(r) for _ in range(4))
            T00, T01 = np.zeros(3), r.normal(size=3) * 0.2
            T10 = np.array([3.2, 0, 0]) + r.normal(size=3) * 0.2
            T11 = np.array([0.2, 0, 0]) + r.normal(size=3) * 0.6
            res = O.query_toc(m1, m2, (R00, T00), (R01, T01), (R10, T10), (R11, T11), 1e-3, 1e-4)
            if res.collisionfree:
                continue
            mo1, mo2 = O.OMotion(R00, T00, R01, T01), O.OMotion(R10, T10, R11, T11)

            def dist(t):
                Ra, Ta, _ = mo1.integrate(t)
                Rb, Tb, _ = mo2.integrate(t)
                return O.distance(m1, m2, Ra, Ta, Rb, Tb, 0.0, 0.0, seeds=(0, 0)).distance
            ts = np.linspace(0, 1, 65)
            first = next((t for t in ts if dist(t) == 0.0), None)
            if first is None:
                continue
            lo, hi = max(0.0, first - 1 / 64), first
            for _ in range(40):
                mid = 0.5 * (lo + hi)
                lo, hi = (lo, mid) if dist(mid) == 0.0 else (mid, hi)
            assert res.toc <= lo, (k, res.toc, lo)
            n_hit += 1
    assert n_hit >= 3


@pytest.mark.parametrize("K", [1, 3, 5, 6])
def test_kernel_logic_strict(emu, K):   # noqa: F811
    """The kernel's strict switches (world axis in the motions it is given,
    leaf rule in the walk and the scan) equal the strict oracle bit for bit."""
    tris, _ = W.torus_knot(n_seg=128, n_ring=12)
    t9 = tris.reshape(-1, 9)
    nq = 24
    P = W.knot_queries(nq, 700 + K, collide_frac=0.4)
    got = run_emu(emu, t9, t9, P, G.default_params(tol_d=1e-3, tol_t=1e-4, strict=1), K)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    with O.strict_mode(True):
        ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                          with_contacts=True, nthreads=4)
    same(ref, got)
    with O.strict_mode(False):
        plain = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                            with_contacts=True, nthreads=4)
    assert not np.array_equal(plain["toc"], got["toc"])   # the switch is live
