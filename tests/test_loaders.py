"""Readers of the demo's .tri / .ani files (c2a_amd/loaders.py) and the
headless replay of the demo's C2A_Solve loop (CCDDemo/mainTorusknot.cpp:
216-321).  The files are written here from known arrays, so the tests need
nothing outside the repository."""
import numpy as np
import pytest

from c2a_amd import loaders as L
from c2a_amd import workloads as W


def write_tri(path, v, f):
    with open(path, "w") as o:
        o.write("TRI\n\n%d\n%d\n" % (len(v), len(f)))
        for x in v:
            o.write("%r %r %r\n" % tuple(float(c) for c in x))
        for t in f:
            o.write("%d %d %d\n" % tuple(int(i) for i in t))


def write_ani(path, R, T):
    with open(path, "w") as o:
        o.write("%df\n" % len(R))
        for i in range(len(R)):
            cols = R[i].T.reshape(9)            # column by column
            o.write("%df\n%s\n%s\n" % (i, " ".join(repr(float(c)) for c in cols),
                                       " ".join(repr(float(c)) for c in T[i])))


def test_tri_small_file(tmp_path):
    p = tmp_path / "quad.tri"
    p.write_text("TRI\n4\n2\n0 0 0\n1 0 0\n1 1 0\n0 1 0.5\n0 1 2\n0 2 3\n")
    v, f = L.read_tri(p)
    assert v.dtype == np.float64 and f.dtype == np.int32
    assert np.array_equal(v, [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.5]])
    assert np.array_equal(f, [[0, 1, 2], [0, 2, 3]])
    s = L.tri_soup(v, f)
    assert s.shape == (2, 9)
    assert np.array_equal(s[1], [0, 0, 0, 1, 1, 0, 0, 1, 0.5])


def test_tri_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.tri"
    p.write_text("OFF\n3\n1\n0 0 0 1 0 0 0 1 0\n0 1 2\n")
    with pytest.raises(ValueError):
        L.read_tri(p)
    p.write_text("TRI\n3\n1\n0 0 0 1 0 0 0 1 0\n0 1 3\n")
    with pytest.raises(ValueError):
        L.read_tri(p)
    p.write_text("TRI\n3\n2\n0 0 0 1 0 0 0 1 0\n0 1 2\n")
    with pytest.raises(ValueError):
        L.read_tri(p)


def test_ani_rotation_is_read_column_by_column(tmp_path):
    """A rotation that is not symmetric tells rows from columns: the file
    holds R[0][0] R[1][0] R[2][0] R[0][1] ... (mainTorusknot.cpp:580-583)."""
    p = tmp_path / "a.ani"
    p.write_text("2f\n0f\n1 0 0 0 1 0 0 0 1\n-300 0 0\n"
                 "1f\n0 1 0 -1 0 0 0 0 1\n1.5 2 3\n")
    R, T = L.read_ani(p)
    assert np.array_equal(R[0], np.eye(3)) and np.array_equal(T[0], [-300, 0, 0])
    # first column (0, 1, 0): x maps to y, a +90 degree turn about z
    assert np.array_equal(R[1], [[0, -1, 0], [1, 0, 0], [0, 0, 1]])
    assert np.array_equal(R[1] @ [1, 0, 0], [0, 1, 0])
    assert np.array_equal(T[1], [1.5, 2, 3])


def test_text_round_trip_of_the_demo_data(tmp_path):
    """The committed bunny and animations, written out in the demo's formats
    and read back, come back bit-identical, and the demo's queries built from
    them are the workload's."""
    d = np.load(W._FIXTURES + "/bunny_noholes.npz")
    write_tri(tmp_path / "bunny.tri", d["vertices"], d["faces"])
    v, f = L.read_tri(tmp_path / "bunny.tri")
    assert np.array_equal(v, d["vertices"]) and np.array_equal(f, d["faces"])
    a = np.load(W._FIXTURES + "/torusknot_ani.npz")
    for k in range(2):
        write_ani(tmp_path / ("k%d.ani" % k), a["R"][k], a["T"][k])
    R1, T1 = L.read_ani(tmp_path / "k0.ani")
    R2, T2 = L.read_ani(tmp_path / "k1.ani")
    assert np.array_equal(R1, a["R"][0]) and np.array_equal(T2, a["T"][1])
    P = L.demo_queries(R1, T1, R2, T2)
    for x, y in zip(P, W.demo_queries()):
        assert np.array_equal(x, y)
    assert len(P[0]) == 303


@pytest.mark.gpu
def test_demo_replay_chained_matches_oracle():
    """The headless replay with the demo's chained last_tri protocol equals
    the oracle's C2A_Solve run over the same 303 queries in order."""
    from oracle import oracle as O
    t9, faces = W.bunny()
    P = W.demo_queries()
    print("demo replay: 303 chained queries", flush=True)
    got = L.replay(t9, P, chain=True, faces=faces)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    for i in range(len(got)):
        res, _, _ = O.solve(m1, m2, *[(x[i, :9].reshape(3, 3), x[i, 9:]) for x in P])
        g = got[i]
        assert g["toc"] == res.toc and g["distance"] == res.distance, i
        assert g["numCA"] == res.numCA and g["collisionfree"] == res.collisionfree, i
        assert (g["last_tri"][0], g["last_tri"][1]) == (m1.m.last_tri, m2.m.last_tri), i
        if not res.collisionfree:
            assert g["num_contact"] == res.num_contact, i
