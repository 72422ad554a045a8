"""The hierarchy built on the device (c2a_mesh_create_device, bvh_gpu.hip)
and the on-disk hierarchy cache (c2a_mesh_save / c2a_mesh_load).

The host build is bit-exact with the oracle's restatement of the reference
build (tests/test_capi_host.py).  The device build must give the same nodes,
triangle order, ids and vertex indices, bit for bit; the GPU tests compare
them on meshes from 1 to 262,144 triangles, including the split's fallback
(every centroid on one side), flat and duplicated geometry.

The CPU tests check the two pieces of the device build that are not the
host's code: the node numbering formula (dfs_first_child) against host-built
trees, and the parallel form of the split's in-place partition (a queue tape
resolved by pointer jumping) against the sequential partition on random
inputs.  The cache tests run on the CPU: a saved hierarchy loads back equal;
other triangles, a damaged, truncated or foreign file are refused."""
import os

import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W


def dfs_first_child(depth, first, rights):
    return 1 + 2 * (depth + first - rights)


def walk_ranges(nodes):
    """(node, first, num, depth, rights) of every node of a host-built tree."""
    fc = nodes[:, 26].astype(np.int64)

    def leaves(v):
        if fc[v] < 0:
            return 1
        return leaves(fc[v]) + leaves(fc[v] + 1)

    out = []
    stack = [(0, 0, 0, 0)]
    while stack:
        v, first, depth, rights = stack.pop()
        n = leaves(v) if fc[v] >= 0 else 1
        out.append((v, first, n, depth, rights))
        if fc[v] >= 0:
            nl = leaves(fc[v])
            stack.append((fc[v] + 1, first + nl, depth + 1, rights + 1))
            stack.append((fc[v], first, depth + 1, rights))
    return out


@pytest.mark.parametrize("level", [1, 2, 3])
def test_node_numbering_formula(level):
    t, _ = W.icosphere(level, 0.2, 3)
    m = G.Mesh(t.reshape(-1, 9))
    nodes, _ = m.export()
    fc = nodes[:, 26].astype(np.int64)
    for v, first, num, depth, rights in walk_ranges(nodes):
        assert nodes[v, 27] == depth
        if num == 1:
            assert fc[v] == -(first + 1)
        else:
            assert fc[v] == dfs_first_child(depth, first, rights), (v, first, depth, rights)


def lomuto(small):
    arr = list(range(len(small)))
    c1 = 0
    for i in range(len(small)):
        if small[arr[i]]:
            arr[i], arr[c1] = arr[c1], arr[i]
            c1 += 1
    return arr


def tape_partition(small):
    """The device build's form: smalls keep their order; the bigs are the
    tape entries [cycles, pushes) after pointer jumping."""
    small = np.asarray(small, bool)
    big = ~small
    bb = np.concatenate([[0], np.cumsum(big)[:-1]])          # bigs before i
    cyc = small & (bb > 0)
    cb = np.concatenate([[0], np.cumsum(cyc)[:-1]])          # cycles before i
    T = int(big.sum() + cyc.sum())
    tape = np.zeros(T, np.int64)
    for i in range(len(small)):
        if big[i]:
            tape[bb[i] + cb[i]] = i
        elif cyc[i]:
            tape[bb[i] + cb[i]] = -(cb[i] + 1)
    while (tape < 0).any():   # pointer jumping
        ref = tape < 0
        tape[ref] = tape[-tape[ref] - 1]
    return list(np.nonzero(small)[0]) + list(tape[int(cyc.sum()):])


def test_partition_by_tape_equals_in_place_partition():
    rng = np.random.default_rng(0)
    for n in list(range(1, 12)) + [37, 64, 200, 1000]:
        for p in (0.0, 0.1, 0.5, 0.9, 1.0):
            for _ in range(6 if n < 100 else 2):
                small = rng.random(n) < p
                assert tape_partition(small) == lomuto(small), (n, p)
    # the long reference chain of a single early "above" triangle
    small = np.ones(500, bool)
    small[0] = False
    assert tape_partition(small) == lomuto(small)


def _cache_mesh():
    t, idx = W.icosphere(2, 0.1, 9)
    return t.reshape(-1, 9), idx


def test_cache_round_trip(tmp_path):
    t9, idx = _cache_mesh()
    path = str(tmp_path / "sphere.c2abvh")
    m = G.Mesh(t9, vidx=idx, cache=path)
    assert not m.from_cache and os.path.exists(path)
    m2 = G.Mesh(t9, vidx=idx, cache=path)
    assert m2.from_cache
    a, b = m.export(), m2.export()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(m.tri_ids(), m2.tri_ids())
    assert m2.max_depth == m.max_depth and m2.nbv == m.nbv


def test_cache_refuses_other_inputs(tmp_path):
    t9, idx = _cache_mesh()
    path = str(tmp_path / "sphere.c2abvh")
    G.Mesh(t9, vidx=idx).save(path)
    moved = t9.copy()
    moved[5, 2] += 1e-12
    for args in [(moved, None, idx), (t9, None, None), (t9, np.arange(len(t9)) + 1, idx),
                 (t9[:-1], None, idx[:-3])]:
        m = G.Mesh(args[0], ids=args[1], vidx=args[2], cache=path)
        assert not m.from_cache   # rebuilt (and the file rewritten for these inputs)
        G.Mesh(t9, vidx=idx).save(path)


def test_cache_refuses_damaged_files(tmp_path):
    t9, idx = _cache_mesh()
    path = tmp_path / "sphere.c2abvh"
    G.Mesh(t9, vidx=idx).save(str(path))
    good = path.read_bytes()
    cases = {
        "flipped": good[:1000] + bytes([good[1000] ^ 1]) + good[1001:],
        "truncated": good[:-8],
        "trailing": good + b"\0",
        "foreign": b"not a hierarchy" + good[15:],
        "empty": b"",
    }
    import ctypes as C
    for name, data in cases.items():
        path.write_bytes(data)
        h = C.c_void_p()
        rc = G.lib().c2a_mesh_load(os.fsencode(str(path)), None, None, None, 0, C.byref(h))
        assert rc == G.ERR_CACHE, name
        assert not h.value
    path.write_bytes(good)
    h = C.c_void_p()
    assert G.lib().c2a_mesh_load(os.fsencode(str(path)), None, None, None, 0, C.byref(h)) == G.OK
    G.lib().c2a_mesh_destroy(h)


# ---- GPU: device build against the host build ------------------------------

def _meshes():
    out = {}
    one = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float64)
    out["one triangle"] = (one, None)
    out["two triangles"] = (np.concatenate([one, one + 0.5]), None)
    t, idx = W.icosphere(1, 0.1, 2)
    out["icosphere 80"] = (t.reshape(-1, 9), idx)
    rng = np.random.default_rng(4)
    soup = rng.normal(size=(65, 9))
    out["soup 65"] = (soup, None)
    out["soup 1000"] = (rng.normal(size=(1000, 9)), None)
    out["identical 300"] = (np.repeat(soup[:1], 300, axis=0), None)   # the half-split fallback
    k = 41   # a flat k x k vertex grid, two triangles per quad (zero-thickness slabs)
    g = np.array([(i * k + j, (i + 1) * k + j, (i + 1) * k + j + 1) for i in range(k - 1)
                  for j in range(k - 1)] +
                 [(i * k + j, (i + 1) * k + j + 1, i * k + j + 1) for i in range(k - 1)
                  for j in range(k - 1)], np.int32)
    xy = np.stack(np.meshgrid(np.linspace(0, 1, k), np.linspace(0, 1, k), indexing="ij"), -1)
    verts = np.concatenate([xy.reshape(-1, 2), np.zeros((k * k, 1))], 1)
    out["flat grid"] = (verts[g].reshape(-1, 9), g)
    t, idx = W.torus_knot()
    out["knot 8192"] = (t.reshape(-1, 9), idx)
    t, idx = W.icosphere(6, 0.05, 5)
    out["icosphere 81920"] = (t.reshape(-1, 9), idx)
    s, idx = W.uv_sphere(512, 256, 0.0, 3)
    out["uv sphere 262144"] = (s.reshape(-1, 9), idx)
    return out


@pytest.mark.gpu
def test_device_build_equals_host_build():
    import time
    for name, (t9, idx) in _meshes().items():
        print("device build: %s" % name, flush=True)
        ids = np.arange(len(t9), dtype=np.int32)[::-1].copy()
        t0 = time.time()
        h = G.Mesh(t9, ids=ids, vidx=idx)
        th = time.time() - t0
        t0 = time.time()
        d = G.Mesh(t9, ids=ids, vidx=idx, device=True)
        td = time.time() - t0
        print("  host %.1f ms, device %.1f ms" % (1e3 * th, 1e3 * td), flush=True)
        hn, ht = h.export()
        dn, dt = d.export()
        assert d.max_depth == h.max_depth, name
        bad = np.nonzero((hn != dn).any(1))[0]
        assert len(bad) == 0, (name, len(bad), bad[:5])
        assert np.array_equal(ht, dt), name
        assert np.array_equal(h.tri_ids(), d.tri_ids()), name


@pytest.mark.gpu
def test_device_built_mesh_runs_queries(knot_tris):
    """Queries on a device-built mesh (its device copy kept from the build)
    give the host-built mesh's results."""
    t9, idx = knot_tris
    P = W.knot_batch_queries(256)
    q = G.make_queries(*P)
    p = G.default_params(tol_d=1e-3, tol_t=1e-4)
    a = G.ccd_batch([G.Mesh(t9, vidx=idx)] * 2, q, p)
    dm = G.Mesh(t9, vidx=idx, device=True)
    b = G.ccd_batch([dm, dm], q, p)
    for f in ("toc", "distance", "numCA", "bv_tests", "tri_tests", "last_tri"):
        assert np.array_equal(a[f], b[f]), f
