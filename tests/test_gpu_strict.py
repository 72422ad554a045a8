"""Strict (conservative) mode on the GPU, through the C ABI.

* Bit for bit against the strict oracle (knot pair, uneven mesh table).
* Zero non-conservative results on the full knot (4096 queries), uneven
  (16,384) and grazing (1024) batches, for the plain order and for the
  learned order with the shipped scorer: no colliding query ends with its
  TOC pose in contact (the final step's distance is exact there), and on the
  knot batch the exact distance (oracle, zero error bounds) is positive at
  eight poses sampled up to each TOC.
* The learned order on the uneven workload (the learned-order evaluation
  workload) against the plain order: fewer BV tests, identical flags.
"""
import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import learn as Lm
from c2a_amd import workloads as W
from oracle import oracle as O
from tests.test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def in_contact_at_toc(r):
    """Colliding queries whose TOC pose is in contact (toc > 0, exact final
    distance 0): the CA loop stepped past the first contact."""
    return np.nonzero((r["collisionfree"] == 0) & (r["toc"] > 0) & (r["distance"] == 0))[0]


@pytest.fixture(scope="module")
def knot():
    tris, idx = W.torus_knot()
    t9 = tris.reshape(-1, 9)
    return t9, [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)], O.OModel(t9), O.OModel(t9)


@pytest.fixture(scope="module")
def scorer():
    return G.Scorer(path=Lm.DEFAULT_SCORER)


def test_strict_knot_bit_exact(knot):
    t9, meshes, m1, m2 = knot
    nq = 256
    print("strict knot batch nq=%d" % nq, flush=True)
    P = W.knot_queries(nq, 515, collide_frac=0.4)
    got = G.ccd_batch(meshes, G.make_queries(*P),
                      G.default_params(tol_d=1e-3, tol_t=1e-4, strict=1))
    with O.strict_mode(True):
        ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                          with_contacts=True, nthreads=8)
    assert_same(ref, got, "strict")


def test_strict_uneven_bit_exact():
    tab = W.uneven_meshes()
    sizes = [len(t) for t, _, _ in tab]
    nq = 256
    P00, P01, P10, P11, mp = W.uneven_queries(nq, 911, sizes, [r for _, _, r in tab])
    meshes = [G.Mesh(t, vidx=idx) for t, idx, _ in tab]
    print("strict uneven nq=%d" % nq, flush=True)
    got = G.ccd_batch(meshes, G.make_queries(P00, P01, P10, P11, mesh_pairs=mp),
                      G.default_params(tol_d=1e-3, tol_t=1e-4, strict=1))
    models = [O.OModel(t) for t, _, _ in tab]
    ref = None
    with O.strict_mode(True):
        for a, b in sorted(set(map(tuple, mp))):
            sel = np.nonzero((mp[:, 0] == a) & (mp[:, 1] == b))[0]
            r = O.batch_toc(models[a], models[b], P00[sel], P01[sel], P10[sel], P11[sel], 1e-3,
                            1e-4, seeds=np.zeros((len(sel), 2), np.int32), with_contacts=True,
                            nthreads=8)
            if ref is None:
                ref = np.zeros(nq, dtype=r.dtype)
            ref[sel] = r
    assert_same(ref, got, "strict uneven")


def _sampled_contact_free(m1, m2, P, res, idx, nsamp=8):
    """Exact distance (oracle, zero error bounds) at nsamp + 1 poses on
    [0, toc] of each listed query; returns the queries touching at one."""
    rows, owner = [], []
    for i in idx:
        mo1 = O.OMotion(P[0][i, :9].reshape(3, 3), P[0][i, 9:], P[1][i, :9].reshape(3, 3),
                        P[1][i, 9:])
        mo2 = O.OMotion(P[2][i, :9].reshape(3, 3), P[2][i, 9:], P[3][i, :9].reshape(3, 3),
                        P[3][i, 9:])
        for t in np.linspace(0.0, res["toc"][i], nsamp + 1):
            Ra, Ta, _ = mo1.integrate(t)
            Rb, Tb, _ = mo2.integrate(t)
            rows.append(np.concatenate([Ra.ravel(), Ta, Rb.ravel(), Tb]))
            owner.append(i)
    if not rows:
        return []
    d = O.distance_batch(m1, m2, np.array(rows), nthreads=16)
    return sorted({owner[k] for k, r in enumerate(d) if r.distance == 0.0})


@pytest.mark.parametrize("order", ["plain", "learned"])
def test_strict_knot_4096_never_past_contact(knot, scorer, order):
    t9, meshes, m1, m2 = knot
    cfg = W.CONFIG_KNOT_BATCH
    P = W.knot_batch_queries()
    print("strict knot 4096 order=%s" % order, flush=True)
    kw = dict(order=G.ORDER_LEARNED, scorer=scorer) if order == "learned" else {}
    res = G.ccd_batch(meshes, G.make_queries(*P),
                      G.default_params(tol_d=cfg["tol_d"], tol_t=cfg["tol_t"], with_contacts=0,
                                       strict=1, **kw))
    assert len(in_contact_at_toc(res)) == 0, in_contact_at_toc(res)[:10]
    coll = np.nonzero((res["collisionfree"] == 0) & (res["toc"] > 0))[0]
    assert len(coll) > 1000
    bad = _sampled_contact_free(m1, m2, P, res, coll)
    assert not bad, bad[:10]


def test_strict_old_knot_stream_never_past_contact(knot):
    """The first knot stream (knot_queries(4096, 20261, 0.3)), on which the
    reference semantics reach contact on 4 queries: strict mode on none."""
    t9, meshes, m1, m2 = knot
    P = W.knot_queries(4096, W.CONFIG_KNOT_BATCH["seed"], 0.3)
    print("strict knot stream 20261", flush=True)
    ref = G.ccd_batch(meshes, G.make_queries(*P),
                      G.default_params(tol_d=1e-3, tol_t=1e-4, with_contacts=0))
    res = G.ccd_batch(meshes, G.make_queries(*P),
                      G.default_params(tol_d=1e-3, tol_t=1e-4, with_contacts=0, strict=1))
    assert len(in_contact_at_toc(ref)) >= 1
    assert len(in_contact_at_toc(res)) == 0, in_contact_at_toc(res)[:10]
    coll = np.nonzero((res["collisionfree"] == 0) & (res["toc"] > 0))[0]
    bad = _sampled_contact_free(m1, m2, P, res, coll)
    assert not bad, bad[:10]


@pytest.mark.parametrize("order", ["plain", "learned"])
def test_strict_uneven_16384_never_past_contact(scorer, order):
    tab = W.uneven_meshes()
    sizes = [len(t) for t, _, _ in tab]
    cfg = W.CONFIG_UNEVEN
    P00, P01, P10, P11, mp = W.uneven_queries(cfg["nq"], cfg["seed"], sizes,
                                              [r for _, _, r in tab])
    meshes = [G.Mesh(t, vidx=idx) for t, idx, _ in tab]
    print("strict uneven 16384 order=%s" % order, flush=True)
    kw = dict(order=G.ORDER_LEARNED, scorer=scorer) if order == "learned" else {}
    res = G.ccd_batch(meshes, G.make_queries(P00, P01, P10, P11, mesh_pairs=mp),
                      G.default_params(tol_d=cfg["tol_d"], tol_t=cfg["tol_t"], with_contacts=0,
                                       strict=1, **kw))
    assert (res["collisionfree"] == 0).sum() > 4000
    assert len(in_contact_at_toc(res)) == 0, in_contact_at_toc(res)[:10]


@pytest.mark.parametrize("order", ["plain", "learned"])
def test_strict_grazing_1024_never_past_contact(scorer, order):
    cfg = W.CONFIG_GRAZING
    s, idx = W.uv_sphere(cfg["n_lon"], cfg["n_lat"], cfg["noise"], 3)
    t9 = s.reshape(-1, 9)
    P = W.grazing_queries(cfg["nq"], cfg["seed"])
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    print("strict grazing 1024 order=%s" % order, flush=True)
    kw = dict(order=G.ORDER_LEARNED, scorer=scorer) if order == "learned" else {}
    res = G.ccd_batch(meshes, G.make_queries(*P),
                      G.default_params(tol_d=cfg["tol_d"], tol_t=cfg["tol_t"], with_contacts=0,
                                       strict=1, **kw))
    assert (res["collisionfree"] == 0).sum() > 300
    assert len(in_contact_at_toc(res)) == 0, in_contact_at_toc(res)[:10]


def test_learned_order_on_uneven_workload(scorer):
    """The learned-order evaluation workload (the 16,384-query uneven batch,
    first 4096 queries): the shipped scorer visits fewer BV tests than the
    plain order with identical collision flags, in both modes."""
    tab = W.uneven_meshes()
    sizes = [len(t) for t, _, _ in tab]
    cfg = W.CONFIG_UNEVEN
    P00, P01, P10, P11, mp = W.uneven_queries(cfg["nq"], cfg["seed"], sizes,
                                              [r for _, _, r in tab])
    n = 4096
    q = G.make_queries(P00[:n], P01[:n], P10[:n], P11[:n], mesh_pairs=mp[:n])
    meshes = [G.Mesh(t, vidx=idx) for t, idx, _ in tab]
    for strict in (0, 1):
        base = dict(tol_d=cfg["tol_d"], tol_t=cfg["tol_t"], with_contacts=0, strict=strict)
        plain = G.ccd_batch(meshes, q, G.default_params(**base))
        learned = G.ccd_batch(meshes, q, G.default_params(order=G.ORDER_LEARNED, scorer=scorer,
                                                          **base))
        c = Lm.compare(plain, learned)
        print("uneven %d strict=%d: %s" % (n, strict, c), flush=True)
        assert np.array_equal(plain["collisionfree"], learned["collisionfree"])
        assert c["bv_ratio"] > 1.3
