"""Readers for the reference demo's data files and a headless replay of the
demo's CCD loop.

  .tri  "TRI", the vertex count, the triangle count, the vertices (x y z),
        then the triangles (three vertex indices): the demo's reader
        (CCDDemo/mainTorusknot.cpp:448-490) and Model::Model
        (CCDDemo/model.cpp:50-124).  The triangle order of the file is the
        order of the AddTri calls, triangle i getting caller id i.
  .ani  the frame count ("303f"), then per frame its number ("0f"), nine
        rotation entries stored column by column, and three translations
        (readXFormFrames, CCDDemo/mainTorusknot.cpp:565-598).

Values are parsed as FP64 from the text, as the demo's fscanf("%lf") and
operator>> into PQP_REAL do.

    python -m c2a_amd.loaders bunny_noholes.tri torusknot1.ani torusknot2.ani
"""
import argparse
import time

import numpy as np


def read_tri(path):
    """(vertices (nv, 3) float64, faces (nt, 3) int32)."""
    with open(path) as f:
        tok = f.read().split()
    if not tok or tok[0] != "TRI":
        raise ValueError("%s: not a TRI file" % path)
    nv, nt = int(tok[1]), int(tok[2])
    if len(tok) < 3 + 3 * nv + 3 * nt:
        raise ValueError("%s: %d vertices and %d triangles announced, file too short"
                         % (path, nv, nt))
    v = np.array(tok[3:3 + 3 * nv], dtype=np.float64).reshape(nv, 3)
    f = np.array(tok[3 + 3 * nv:3 + 3 * nv + 3 * nt], dtype=np.int32).reshape(nt, 3)
    if nt and (f.min() < 0 or f.max() >= nv):
        raise ValueError("%s: vertex index out of range" % path)
    return v, f


def tri_soup(vertices, faces):
    """(nt, 9) float64: p1, p2, p3 of each triangle, the demo's AddTri arguments."""
    return np.ascontiguousarray(vertices[faces].reshape(-1, 9))


def read_ani(path):
    """(R (nf, 3, 3), T (nf, 3)) per frame; frames absent from the file stay
    zero, as in the demo's static frame table."""
    with open(path) as f:
        tok = f.read().split()
    nf = int(float(tok[0].rstrip("f")))
    R = np.zeros((nf, 3, 3))
    T = np.zeros((nf, 3))
    k = 1
    while k + 12 < len(tok):
        i = int(float(tok[k].rstrip("f")))
        if not 0 <= i < nf:
            raise ValueError("%s: frame %d outside 0..%d" % (path, i, nf - 1))
        vals = np.array(tok[k + 1:k + 13], dtype=np.float64)
        R[i] = vals[:9].reshape(3, 3).T   # stored column by column
        T[i] = vals[9:]
        k += 13
    return R, T


def demo_queries(R1, T1, R2, T2):
    """The demo's C2A_Solve queries, one per frame of animation 2
    (CCDDemo/mainTorusknot.cpp:216-266): object 1 moves from frame step1 to
    step2 of animation 1 ((0, 1) for frames below 101, (101, 102) below 202,
    (202, 203) after); object 2 keeps the rotation of the current frame of
    animation 2 and moves between the translations of frames step1 and step2.
    Returns (P00, P01, P10, P11), each (n, 12) = R row-major then T."""
    n = R2.shape[0]
    P = [np.zeros((n, 12)) for _ in range(4)]
    for f in range(n):
        s1, s2 = (0, 1) if f < 101 else ((101, 102) if f < 202 else (202, 203))
        P[0][f] = np.concatenate([R1[s1].reshape(9), T1[s1]])
        P[1][f] = np.concatenate([R1[s2].reshape(9), T1[s2]])
        P[2][f] = np.concatenate([R2[f].reshape(9), T2[s1]])
        P[3][f] = np.concatenate([R2[f].reshape(9), T2[s2]])
    return P


def replay(tris, P, chain=True, params=None, faces=None):
    """Runs the demo's queries on the GPU with C2A_Solve's parameters.

    chain=True follows the demo's protocol (mainTorusknot.cpp:312-321): each
    query is seeded with the models' last_tri as the previous query left it,
    so the queries run one after another.  chain=False runs them as one batch,
    every query seeded with triangle 0.  Returns the result array."""
    import c2a_amd as G
    t9 = np.ascontiguousarray(tris.reshape(-1, 9))
    meshes = [G.Mesh(t9, vidx=faces), G.Mesh(t9, vidx=faces)]
    p = params or G.default_params()
    if not chain:
        return G.ccd_batch(meshes, G.make_queries(*P), p)
    n = len(P[0])
    out = np.zeros(n, dtype=G.RESULT_DTYPE)
    seed = np.zeros((1, 2), np.int32)
    for i in range(n):
        q = G.make_queries(*[x[i:i + 1] for x in P], seeds=seed)
        r = G.ccd_batch(meshes, q, p)
        out[i] = r[0]
        seed = r["last_tri"].astype(np.int32).reshape(1, 2)
    return out


def main():
    ap = argparse.ArgumentParser(description="headless replay of the reference demo")
    ap.add_argument("tri")
    ap.add_argument("ani1")
    ap.add_argument("ani2")
    ap.add_argument("--batch", action="store_true",
                    help="one batch, every query seeded with triangle 0 (default: the "
                         "demo's chained last_tri protocol)")
    a = ap.parse_args()
    v, f = read_tri(a.tri)
    R1, T1 = read_ani(a.ani1)
    R2, T2 = read_ani(a.ani2)
    P = demo_queries(R1, T1, R2, T2)
    t0 = time.perf_counter()
    r = replay(tri_soup(v, f), P, chain=not a.batch, faces=f)
    dt = time.perf_counter() - t0
    hit = r["collisionfree"] == 0
    print("%d triangles, %d queries in %.3f s: %d collisions, mean toc of those %.6f, "
          "%d CA iterations" % (len(f), len(r), dt, int(hit.sum()),
                                float(r["toc"][hit].mean()) if hit.any() else 0.0,
                                int(r["numCA"].sum())))


if __name__ == "__main__":
    main()
