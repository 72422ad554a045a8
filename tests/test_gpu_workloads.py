"""GPU parity on the other workloads: a batch over a table of meshes of very
different sizes (icospheres of 80 to 81,920 triangles, a torus, the knot) and
near-tangential passes between two 262,144-triangle noisy spheres.  Every
field is compared with the oracle for equality."""
import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = [("toc", "toc"), ("distance", "distance"), ("mint", "mint"), ("numCA", "numCA"),
          ("collisionfree", "collisionfree"), ("num_bv_tests", "bv_tests"),
          ("num_tri_tests", "tri_tests"), ("num_contact", "num_contact"), ("flags", "flags")]


def check(ref, got, label):
    for f, g in FIELDS:
        bad = np.nonzero(ref[f] != got[g])[0]
        assert len(bad) == 0, "%s: %s differs at %s" % (label, f, bad[:5])
    assert np.array_equal(ref["out_last_triA"], got["last_tri"][:, 0]), label
    assert np.array_equal(ref["out_last_triB"], got["last_tri"][:, 1]), label


@pytest.mark.parametrize("nq", [384, W.CONFIG_UNEVEN["nq"]])
def test_uneven_mesh_table_bit_exact(nq):
    """The first 384 queries, then the whole 16,384-query batch."""
    tab = W.uneven_meshes()
    sizes = [len(t) for t, _, _ in tab]
    P00, P01, P10, P11, mp = W.uneven_queries(nq, W.CONFIG_UNEVEN["seed"], sizes,
                                              [r for _, _, r in tab])
    meshes = [G.Mesh(t, vidx=idx) for t, idx, _ in tab]
    print("uneven: %d queries over %d meshes (%s triangles)" % (nq, len(tab), sizes), flush=True)
    got = G.ccd_batch(meshes, G.make_queries(P00, P01, P10, P11, mesh_pairs=mp),
                      G.default_params(tol_d=1e-3, tol_t=1e-4))
    models = [O.OModel(t) for t, _, _ in tab]
    refs = None
    for a, b in sorted(set(map(tuple, mp))):
        sel = np.nonzero((mp[:, 0] == a) & (mp[:, 1] == b))[0]
        r = O.batch_toc(models[a], models[b], P00[sel], P01[sel], P10[sel], P11[sel], 1e-3, 1e-4,
                        seeds=np.zeros((len(sel), 2), np.int32), with_contacts=True, nthreads=16)
        if refs is None:
            refs = np.zeros(nq, dtype=r.dtype)
        refs[sel] = r
    check(refs, got, "uneven")
    assert len(set(map(tuple, mp))) > 20
    assert (got["collisionfree"] == 0).any() and (got["collisionfree"] == 1).any()


@pytest.mark.parametrize("nq", [64, W.CONFIG_GRAZING["nq"]])
def test_grazing_262k_bit_exact(nq):
    """The first 64 passes, then the whole 1024-query batch."""
    cfg = W.CONFIG_GRAZING
    s, idx = W.uv_sphere(cfg["n_lon"], cfg["n_lat"], cfg["noise"], 3)
    t9 = s.reshape(-1, 9)
    P = W.grazing_queries(nq, cfg["seed"])
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    print("grazing: %d queries, %d triangles per sphere" % (nq, len(t9)), flush=True)
    got = G.ccd_batch(meshes, G.make_queries(*P), G.default_params(tol_d=cfg["tol_d"],
                                                                    tol_t=cfg["tol_t"]))
    m1, m2 = O.OModel(t9), O.OModel(t9)
    ref = O.batch_toc(m1, m2, *P, cfg["tol_d"], cfg["tol_t"], seeds=np.zeros((nq, 2), np.int32),
                      with_contacts=True, nthreads=16)
    check(ref, got, "grazing")


def test_reference_demo_batch_bit_exact():
    """The reference demo's 303 bunny-vs-bunny queries as one batch
    (independent queries, C2A_Solve's tolerances and contact count)."""
    t9, idx = W.bunny()
    P = W.demo_queries()
    nq = len(P[0])
    print("demo batch nq=%d, 69,664 triangles per bunny" % nq, flush=True)
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    got = G.ccd_batch(meshes, G.make_queries(*P), G.default_params())
    ref = O.batch_toc(O.OModel(t9), O.OModel(t9), *P, 1e-4, 1e-4,
                      seeds=np.zeros((nq, 2), np.int32), with_contacts=True, nthreads=8)
    check(ref, got, "demo")
    assert (got["collisionfree"] == 0).sum() > 100
