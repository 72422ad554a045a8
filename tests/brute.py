"""Independent brute-force geometry used to pin the oracle (numpy, no shared code).

Closest distance between two convex planar polygons (rectangles, triangles)
in 3-D is attained between an edge of one and an edge of the other, or
between a vertex of one and the face of the other, unless they intersect.
Everything here is the textbook construction of those candidates.
"""
import numpy as np


def seg_seg(p0, p1, q0, q1):
    """Distance between segments [p0,p1] and [q0,q1] (clamped closed form).
    Candidates: interior-interior stationary point (if inside) and the four
    endpoint-to-segment distances."""
    d1, d2, r = p1 - p0, q1 - q0, p0 - q0
    best = min(pt_seg(p0, q0, q1), pt_seg(p1, q0, q1), pt_seg(q0, p0, p1), pt_seg(q1, p0, p1))
    a, e, b = d1 @ d1, d2 @ d2, d1 @ d2
    c, f = d1 @ r, d2 @ r
    den = a * e - b * b
    if den > 1e-300 * max(a * e, 1e-300):
        s = (b * f - c * e) / den
        t = (a * f - b * c) / den
        if 0 <= s <= 1 and 0 <= t <= 1:
            best = min(best, np.linalg.norm((p0 + s * d1) - (q0 + t * d2)))
    return best


def pt_seg(p, a, b):
    ab = b - a
    den = ab @ ab
    t = 0.0 if den == 0 else np.clip((p - a) @ ab / den, 0.0, 1.0)
    return float(np.linalg.norm(p - (a + t * ab)))


def pt_poly(p, verts):
    """Distance from p to a planar convex polygon (vertices in order)."""
    n = np.cross(verts[1] - verts[0], verts[2] - verts[0])
    nn = np.linalg.norm(n)
    best = min(pt_seg(p, verts[i], verts[(i + 1) % len(verts)]) for i in range(len(verts)))
    if nn > 0:
        n = n / nn
        h = (p - verts[0]) @ n
        proj = p - h * n
        inside = all(np.cross(verts[(i + 1) % len(verts)] - verts[i], proj - verts[i]) @ n >= 0
                     for i in range(len(verts)))
        if inside:
            best = min(best, abs(h))
    return best


def seg_crosses_poly(a, b, verts):
    n = np.cross(verts[1] - verts[0], verts[2] - verts[0])
    if np.linalg.norm(n) == 0:
        return False
    da, db = (a - verts[0]) @ n, (b - verts[0]) @ n
    if da * db > 0 or da == db:
        return False
    x = a + (da / (da - db)) * (b - a)
    return all(np.cross(verts[(i + 1) % len(verts)] - verts[i], x - verts[i]) @ n >= 0
               for i in range(len(verts)))


def poly_poly(A, B):
    """Distance between two convex planar polygons (vertex arrays in order)."""
    ea = [(A[i], A[(i + 1) % len(A)]) for i in range(len(A))]
    eb = [(B[i], B[(i + 1) % len(B)]) for i in range(len(B))]
    for a, b in ea:
        if seg_crosses_poly(a, b, B):
            return 0.0
    for a, b in eb:
        if seg_crosses_poly(a, b, A):
            return 0.0
    best = min(seg_seg(a0, a1, b0, b1) for a0, a1 in ea for b0, b1 in eb)
    best = min([best] + [pt_poly(v, B) for v in A] + [pt_poly(v, A) for v in B])
    return best


def rect_corners(l0, l1, R=None, T=None):
    c = np.array([[0, 0, 0], [l0, 0, 0], [l0, l1, 0], [0, l1, 0]], dtype=np.float64)
    if R is not None:
        c = c @ np.asarray(R).T + np.asarray(T)
    return c


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def rodrigues(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def mesh_min_distance(trisA, trisB, RA, TA, RB, TB):
    """All-pairs triangle distance at the given poses (tiny meshes only)."""
    A = trisA.reshape(-1, 3, 3) @ np.asarray(RA).T + np.asarray(TA)
    B = trisB.reshape(-1, 3, 3) @ np.asarray(RB).T + np.asarray(TB)
    # cheap prefilter by bounding spheres
    ca, cb = A.mean(1), B.mean(1)
    ra = np.linalg.norm(A - ca[:, None], axis=2).max(1)
    rb = np.linalg.norm(B - cb[:, None], axis=2).max(1)
    dc = np.linalg.norm(ca[:, None] - cb[None], axis=2) - ra[:, None] - rb[None]
    lb = dc.min()
    best = np.inf
    order = np.argsort(dc, axis=None)
    for k in order:
        i, j = divmod(int(k), B.shape[0])
        if dc[i, j] > best:
            break
        best = min(best, poly_poly(A[i], B[j]))
    return best, lb


def count_pairs_within(trisA, trisB, RA, TA, RB, TB, thr):
    A = trisA.reshape(-1, 3, 3) @ np.asarray(RA).T + np.asarray(TA)
    B = trisB.reshape(-1, 3, 3) @ np.asarray(RB).T + np.asarray(TB)
    ca, cb = A.mean(1), B.mean(1)
    ra = np.linalg.norm(A - ca[:, None], axis=2).max(1)
    rb = np.linalg.norm(B - cb[:, None], axis=2).max(1)
    dc = np.linalg.norm(ca[:, None] - cb[None], axis=2) - ra[:, None] - rb[None]
    n = 0
    for i, j in zip(*np.nonzero(dc <= thr)):
        if poly_poly(A[i], B[j]) <= thr:
            n += 1
    return n


def touching(trisA, trisB, RA, TA, RB, TB):
    """True when some triangle pair intersects (distance 0) at the poses."""
    A = trisA.reshape(-1, 3, 3) @ np.asarray(RA).T + np.asarray(TA)
    B = trisB.reshape(-1, 3, 3) @ np.asarray(RB).T + np.asarray(TB)
    ca, cb = A.mean(1), B.mean(1)
    ra = np.linalg.norm(A - ca[:, None], axis=2).max(1)
    rb = np.linalg.norm(B - cb[:, None], axis=2).max(1)
    dc = np.linalg.norm(ca[:, None] - cb[None], axis=2) - ra[:, None] - rb[None]
    cand = np.argwhere(dc <= 0)
    for i, j in cand[np.argsort(dc[dc <= 0])]:
        if poly_poly(A[i], B[j]) == 0.0:
            return True
    return False


def first_touch_time(trisA, trisB, poseA, poseB, n=100):
    """First sampled t in [0,1] at which the meshes touch (None if never).
    poseA/poseB: callables t -> (R, T)."""
    for t in np.linspace(0.0, 1.0, n + 1):
        RA, TA = poseA(t)
        RB, TB = poseB(t)
        if touching(trisA, trisB, RA, TA, RB, TB):
            return float(t)
    return None
