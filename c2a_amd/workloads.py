"""Seeded synthetic workloads for the CCD engine.

Meshes are triangle soups given as (n, 3, 3) float64 arrays (vertex xyz per
triangle corner) plus an (n, 3) int32 vertex-index array.  Queries are four
poses per query (start/end of object 1 and 2), each a (nq, 12) float64 array
laid out as a row-major 3x3 rotation followed by the translation.

All randomness comes from a counter-based splitmix64 stream so the same seed
gives the same bytes on every machine.  The reference ships no knot mesh
(its demo uses the bunny, CCDDemo/mainTorusknot.cpp:53-54); the shapes below
are this project's choices and are documented with their parameters.
"""
import math
import os

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix(x):
    with np.errstate(over="ignore"):
        z = x + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Rng:
    """Counter-based uniform stream: value i = splitmix64(key + i*golden)."""

    def __init__(self, seed, stream=0):
        with np.errstate(over="ignore"):
            self.key = _splitmix(np.uint64(seed) * np.uint64(1000003) + np.uint64(stream))
        self.ctr = 0

    def uniform(self, n, lo=0.0, hi=1.0):
        idx = np.arange(self.ctr, self.ctr + n, dtype=np.uint64)
        self.ctr += n
        with np.errstate(over="ignore"):
            z = _splitmix(self.key + idx * _GOLDEN)
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u


# ---------------------------------------------------------------------------
# rotations
# ---------------------------------------------------------------------------
def quat_to_matrix(q):
    """Unit quaternions (n, 4) as (x, y, z, w) -> rotation matrices (n, 3, 3)."""
    q = np.asarray(q, dtype=np.float64)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def random_quats(rng, n):
    """Uniform random unit quaternions (Shoemake's subgroup algorithm)."""
    u1, u2, u3 = rng.uniform(n), rng.uniform(n), rng.uniform(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)


def pack_poses(R, T):
    """(n,3,3),(n,3) -> (n,12)."""
    return np.concatenate([R.reshape(-1, 9), T.reshape(-1, 3)], axis=1).astype(np.float64)


def unpack_pose(p12):
    return p12[:9].reshape(3, 3), p12[9:12]


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def _grid_tris(nu, nv, wrap_v=True):
    """Two triangles per quad of a (nu x nv) vertex grid wrapped in u (and v)."""
    tris = []
    for i in range(nu):
        i1 = (i + 1) % nu
        for j in range(nv if wrap_v else nv - 1):
            j1 = (j + 1) % nv
            a, b, c, d = i * nv + j, i1 * nv + j, i1 * nv + j1, i * nv + j1
            tris.append((a, b, c))
            tris.append((a, c, d))
    return np.array(tris, dtype=np.int32)


def torus_knot(p=2, q=3, n_seg=256, n_ring=16, tube=0.1):
    """(p,q) torus-knot tube.  Centre line
    c(phi) = [(2 + cos q*phi) cos p*phi, (2 + cos q*phi) sin p*phi, sin q*phi] / 3
    (maximum radius 1.0), tube radius `tube`, parallel-transport frame with the
    closing twist spread evenly along the knot.  256x16 quads -> 8192 triangles.
    """
    phi = 2 * np.pi * np.arange(n_seg) / n_seg
    c = np.stack([(2 + np.cos(q * phi)) * np.cos(p * phi),
                  (2 + np.cos(q * phi)) * np.sin(p * phi),
                  np.sin(q * phi)], axis=1) / 3.0
    tan = np.roll(c, -1, axis=0) - np.roll(c, 1, axis=0)
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    ref = np.array([0.0, 0.0, 1.0])
    if abs(tan[0] @ ref) > 0.9:
        ref = np.array([1.0, 0.0, 0.0])
    N = np.empty_like(c)
    n = ref - (ref @ tan[0]) * tan[0]
    N[0] = n / np.linalg.norm(n)
    for i in range(1, n_seg):
        n = N[i - 1] - (N[i - 1] @ tan[i]) * tan[i]
        N[i] = n / np.linalg.norm(n)
    # closing twist: angle between the transported frame and N[0] about tan[0]
    n_end = N[-1] - (N[-1] @ tan[0]) * tan[0]
    n_end /= np.linalg.norm(n_end)
    twist = math.atan2(np.cross(n_end, N[0]) @ tan[0], n_end @ N[0])
    B = np.cross(tan, N)
    ang = twist * np.arange(n_seg) / n_seg
    Nr = np.cos(ang)[:, None] * N + np.sin(ang)[:, None] * B
    Br = np.cross(tan, Nr)
    th = 2 * np.pi * np.arange(n_ring) / n_ring
    verts = (c[:, None, :] + tube * (np.cos(th)[None, :, None] * Nr[:, None, :] +
                                     np.sin(th)[None, :, None] * Br[:, None, :]))
    verts = verts.reshape(-1, 3)
    idx = _grid_tris(n_seg, n_ring)
    return verts[idx], idx


def torus(n_major=64, n_minor=32, R=1.0, r=0.3):
    """Ring torus, n_major x n_minor quads -> 2*n_major*n_minor triangles."""
    u = 2 * np.pi * np.arange(n_major) / n_major
    v = 2 * np.pi * np.arange(n_minor) / n_minor
    cu, su = np.cos(u)[:, None], np.sin(u)[:, None]
    cv, sv = np.cos(v)[None, :], np.sin(v)[None, :]
    verts = np.stack([(R + r * cv) * cu, (R + r * cv) * su, r * sv + 0 * cu], axis=2)
    verts = verts.reshape(-1, 3)
    idx = _grid_tris(n_major, n_minor)
    return verts[idx], idx


def _value_noise(rng, pts, freq=3.0, lattice=16):
    """Trilinear value noise in [-1, 1] over a seeded periodic lattice."""
    vals = rng.uniform(lattice ** 3, -1.0, 1.0).reshape(lattice, lattice, lattice)
    x = (pts + 2.0) * freq
    i0 = np.floor(x).astype(np.int64)
    f = x - i0
    f = f * f * (3 - 2 * f)
    out = np.zeros(pts.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) *
                     (f[:, 2] if dz else 1 - f[:, 2]))
                ii = (i0 + np.array([dx, dy, dz])) % lattice
                out += w * vals[ii[:, 0], ii[:, 1], ii[:, 2]]
    return out


def icosphere(level, noise_amp=0.0, seed=0):
    """Icosahedron subdivided `level` times (20*4^level triangles), projected
    to the unit sphere and displaced radially by amp * value noise."""
    t = (1 + math.sqrt(5)) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = np.array(f, dtype=np.int64)
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = np.array(nf, dtype=np.int64)
    V = np.array(verts)
    if noise_amp:
        V = V * (1.0 + noise_amp * _value_noise(Rng(seed, 7), V))[:, None]
    return V[faces], faces.astype(np.int32)


def uv_sphere(n_lon=512, n_lat=256, noise_amp=0.0, seed=0):
    """UV sphere with n_lat latitude bands between n_lat+1 rings placed at
    theta = pi*(i+0.5)/(n_lat+1) (the two polar caps stay open), n_lon
    longitudes: 2*n_lon*n_lat triangles (512x256 -> 262,144)."""
    th = np.pi * (np.arange(n_lat + 1) + 0.5) / (n_lat + 1)
    ph = 2 * np.pi * np.arange(n_lon) / n_lon
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    V = np.stack([st * np.cos(ph)[None, :], st * np.sin(ph)[None, :], ct + 0 * ph[None, :]],
                 axis=2).reshape(-1, 3)
    if noise_amp:
        V = V * (1.0 + noise_amp * _value_noise(Rng(seed, 11), V))[:, None]
    tris = []
    for i in range(n_lat):
        for j in range(n_lon):
            j1 = (j + 1) % n_lon
            a, b, c, d = i * n_lon + j, (i + 1) * n_lon + j, (i + 1) * n_lon + j1, i * n_lon + j1
            tris.append((a, b, c))
            tris.append((a, c, d))
    idx = np.array(tris, dtype=np.int32)
    return V[idx], idx


# ---------------------------------------------------------------------------
# query sets
# ---------------------------------------------------------------------------
def random_poses(rng, n, span=2.0):
    R = quat_to_matrix(random_quats(rng, n))
    T = rng.uniform(3 * n, -span, span).reshape(n, 3)
    return R, T


def knot_queries(nq, seed, collide_frac=0.0):
    """Start/end poses for object 1 and 2: translation U[-2,2]^3, uniform
    random rotation.  With probability collide_frac the end pose of object 2
    is pulled toward object 1's end pose: T <- T + a*(T1end - T), a~U[0.8,1].
    Returns (P00, P01, P10, P11), each (nq, 12)."""
    rng = Rng(seed, 1)
    R00, T00 = random_poses(rng, nq)
    R01, T01 = random_poses(rng, nq)
    R10, T10 = random_poses(rng, nq)
    R11, T11 = random_poses(rng, nq)
    if collide_frac > 0:
        pick = rng.uniform(nq) < collide_frac
        a = rng.uniform(nq, 0.8, 1.0)[:, None]
        T11 = np.where(pick[:, None], T11 + a * (T01 - T11), T11)
    return (pack_poses(R00, T00), pack_poses(R01, T01), pack_poses(R10, T10),
            pack_poses(R11, T11))


CONFIG_KNOT_SOLVE = dict(name="knot-pair sequential C2A_Solve parity", nq=1000, seed=20260,
                         tol_d=1e-3, tol_t=1e-4, collide_frac=0.0)
# The knot batch's queries: knot_batch_queries() (about 30% colliding).  The
# scorer's training data used knot_queries(4096, seed, 0.3) with this seed.
CONFIG_KNOT_BATCH = dict(name="knot-pair batched CCD", nq=4096, seed=20261,
                         tol_d=1e-3, tol_t=1e-4, collide_frac=0.3,
                         fixture="knot_batch_4096.npz")


def knot_batch_queries(nq=None, path=None):
    """The knot batch (4096 queries, about 30% colliding), or its first nq.

    Poses as knot_queries: translation U[-2,2]^3 and uniform rotations, with
    the colliding queries' end pose of object 2 pulled toward object 1's by
    the seeded lerp.  Two radius-1 knots that close collide on most plain
    draws, so the batch takes, per query, either the next colliding pulled
    candidate or the next collision-free plain candidate of two seeded
    streams; which candidates those are was classified once with the FP64
    oracle and is stored in tests/golden/knot_batch_4096.npz (written by
    tests/golden/make_knot_batch.py).  Returns (P00, P01, P10, P11)."""
    d = np.load(path or os.path.join(_FIXTURES, CONFIG_KNOT_BATCH["fixture"]))
    seed_free, seed_hit, _ = (int(v) for v in d["seeds"])
    n_free, n_hit = (int(v) for v in d["n_cand"])
    is_hit, fi, hi = d["is_hit"], d["free_idx"], d["hit_idx"]
    free = knot_queries(n_free, seed_free, 0.0)
    hit = knot_queries(n_hit, seed_hit, 1.0)
    out = []
    for f, h in zip(free, hit):
        P = np.empty((len(is_hit), 12))
        P[~is_hit] = f[fi]
        P[is_hit] = h[hi]
        out.append(P[:nq] if nq else P)
    return tuple(out)


# ---------------------------------------------------------------------------
# uneven batch: many meshes of very different sizes
# ---------------------------------------------------------------------------
def uneven_meshes(levels=(1, 2, 3, 4, 5, 6), seed=7):
    """Mesh table of the uneven batch: noise-displaced icospheres (amplitude
    0.1) at the given subdivision levels (80 * 4^(L-1) triangles: 80 ... 81,920),
    a 64x32 torus and the (2,3) knot tube.  Returns a list of (tris9, vidx,
    bounding radius about the origin)."""
    out = []
    for lv in levels:
        t, idx = icosphere(lv, 0.1, seed + lv)
        out.append((t.reshape(-1, 9), idx))
    t, idx = torus(64, 32)
    out.append((t.reshape(-1, 9), idx))
    t, idx = torus_knot()
    out.append((t.reshape(-1, 9), idx))
    return [(t, idx, float(np.sqrt((t.reshape(-1, 3) ** 2).sum(1)).max())) for t, idx in out]


def _rotation_about(axis, angle):
    axis = axis / np.linalg.norm(axis, axis=1, keepdims=True)
    h = 0.5 * angle
    q = np.concatenate([axis * np.sin(h)[:, None], np.cos(h)[:, None]], axis=1)
    return quat_to_matrix(q)


def uneven_queries(nq, seed, sizes, radii):
    """Queries over a mesh table: each object's mesh is drawn log-uniformly in
    triangle count (nearest table entry in log size), and each object moves
    rigidly: random start rotation, end rotation = start turned about a random
    axis by U[0, pi], translation span U[0.5, 4] body radii along a random
    direction.  Object 2 starts U[1, 3] combined radii from object 1 and its
    path is aimed at object 1's end position with probability 0.5.  Returns
    (P00, P01, P10, P11, mesh_pairs)."""
    rng = Rng(seed, 3)
    sizes = np.asarray(sizes, dtype=np.float64)
    radii = np.asarray(radii, dtype=np.float64)
    ls = np.log(sizes)
    pick = rng.uniform(2 * nq, ls.min(), ls.max()).reshape(nq, 2)
    mp = np.abs(pick[:, :, None] - ls[None, None, :]).argmin(axis=2).astype(np.int32)
    r1, r2 = radii[mp[:, 0]], radii[mp[:, 1]]

    def unit(n):
        v = rng.uniform(3 * n, -1.0, 1.0).reshape(n, 3)
        return v / np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-12)

    def motion(r, start):
        R0 = quat_to_matrix(random_quats(rng, nq))
        R1 = np.einsum("nij,njk->nik", _rotation_about(unit(nq), rng.uniform(nq, 0.0, np.pi)), R0)
        T1 = start + unit(nq) * (rng.uniform(nq, 0.5, 4.0) * r)[:, None]
        return R0, start, R1, T1

    R00, T00, R01, T01 = motion(r1, np.zeros((nq, 3)))
    start2 = unit(nq) * (rng.uniform(nq, 1.0, 3.0) * (r1 + r2))[:, None]
    R10, T10, R11, T11 = motion(r2, start2)
    aim = rng.uniform(nq) < 0.5
    T11 = np.where(aim[:, None], T01 + unit(nq) * (0.5 * r1)[:, None], T11)
    return (pack_poses(R00, T00), pack_poses(R01, T01), pack_poses(R10, T10),
            pack_poses(R11, T11), mp)


# ---------------------------------------------------------------------------
# large grazing contacts
# ---------------------------------------------------------------------------
def grazing_queries(nq, seed, radius=1.0, max_sep=5e-3):
    """Two unit-size bodies, object 1 fixed at the origin turning slowly,
    object 2 passing it on a straight line whose closest approach between the
    centers is 2*radius*(1 + s), s ~ U[0, max_sep] (near-tangential motion),
    while turning about a random axis.  Returns (P00, P01, P10, P11)."""
    rng = Rng(seed, 5)

    def unit(n):
        v = rng.uniform(3 * n, -1.0, 1.0).reshape(n, 3)
        return v / np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-12)

    R00 = quat_to_matrix(random_quats(rng, nq))
    R01 = np.einsum("nij,njk->nik", _rotation_about(unit(nq), rng.uniform(nq, 0.0, 0.2)), R00)
    T0 = np.zeros((nq, 3))
    n = unit(nq)                                  # direction of closest approach
    d = unit(nq)
    d = d - (d * n).sum(1, keepdims=True) * n     # travel direction, orthogonal to n
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    sep = rng.uniform(nq, 0.0, max_sep)
    c = n * (2 * radius * (1 + sep))[:, None]
    half = rng.uniform(nq, 1.0, 2.0)[:, None] * radius
    R10 = quat_to_matrix(random_quats(rng, nq))
    R11 = np.einsum("nij,njk->nik", _rotation_about(unit(nq), rng.uniform(nq, 0.0, 1.0)), R10)
    return (pack_poses(R00, T0), pack_poses(R01, T0), pack_poses(R10, c - half * d),
            pack_poses(R11, c + half * d))


CONFIG_UNEVEN = dict(name="uneven batched CCD", nq=16384, seed=20262, tol_d=1e-3, tol_t=1e-4)
CONFIG_GRAZING = dict(name="large grazing-contact stress", nq=1024, seed=20264, tol_d=1e-3,
                      tol_t=1e-4, n_lon=512, n_lat=256, noise=0.1)


# ---------------------------------------------------------------------------
# The reference demo (CCDDemo/mainTorusknot.cpp): two copies of the Stanford
# bunny (tri_models/bunny_noholes.tri, 69,664 triangles) moved by the two
# torus-knot animation files (models/torusknot{1,2}.ani, 303 frames); the
# data are fixtures written by tests/golden/make_fixtures.py.
# ---------------------------------------------------------------------------
_FIXTURES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                         "golden")


def bunny(path=None):
    """(triangles (n, 9) float64, vertex indices (n, 3) int32) of the demo's
    bunny, in the file's triangle order."""
    d = np.load(path or os.path.join(_FIXTURES, "bunny_noholes.npz"))
    v, f = d["vertices"], d["faces"]
    return v[f].reshape(-1, 9).copy(), f.astype(np.int32)


def demo_queries(path=None):
    """The demo's 303 C2A_Solve queries (CCDDemo/mainTorusknot.cpp:216-266):
    object 1 moves from frame step1 to step2 of animation 1 ((0, 1) for
    frames below 101, (101, 102) below 202, (202, 203) after); object 2
    keeps the rotation of frame `iframe` of animation 2 and moves between
    the translations of frames step1 and step2.  Returns (P00, P01, P10, P11)."""
    from c2a_amd import loaders
    d = np.load(path or os.path.join(_FIXTURES, "torusknot_ani.npz"))
    R, T = d["R"], d["T"]
    return loaders.demo_queries(R[0], T[0], R[1], T[1])


CONFIG_DEMO = dict(name="reference demo: bunny vs bunny, torus-knot animation", nq=303,
                   tol_d=1e-4, tol_t=1e-4)
