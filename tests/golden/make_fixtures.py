"""Writes the reference's own demo data as compact fixtures (run once, in a
container that has the reference tree; the fixtures are committed).

  bunny_noholes.npz  <- tri_models/bunny_noholes.tri   (vertices, faces)
  torusknot_ani.npz  <- models/torusknot1.ani, torusknot2.ani (R, T per frame)

The .tri format is "TRI", the vertex count, the triangle count, the vertices
(x y z) and the triangles (three vertex indices), as the demo reads it
(CCDDemo/mainTorusknot.cpp:448-490).  An .ani file is the frame count
("303f"), then per frame its number ("0f"), nine rotation entries stored
column by column, and three translations (readXFormFrames,
CCDDemo/mainTorusknot.cpp:565-598).  Values are parsed as FP64 from the text,
as the demo's fscanf("%lf") / operator>> do.

    python tests/golden/make_fixtures.py /root/reference
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from c2a_amd.loaders import read_ani, read_tri  # noqa: E402


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    v, f = read_tri(os.path.join(ref, "tri_models", "bunny_noholes.tri"))
    np.savez_compressed(os.path.join(HERE, "bunny_noholes.npz"), vertices=v, faces=f)
    R1, T1 = read_ani(os.path.join(ref, "models", "torusknot1.ani"))
    R2, T2 = read_ani(os.path.join(ref, "models", "torusknot2.ani"))
    np.savez_compressed(os.path.join(HERE, "torusknot_ani.npz"), R=np.stack([R1, R2]),
                        T=np.stack([T1, T2]))
    print("bunny: %d vertices, %d triangles; animation: %d frames" % (len(v), len(f), len(R1)))


if __name__ == "__main__":
    main()
