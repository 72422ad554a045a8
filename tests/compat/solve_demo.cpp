// solve_demo.cpp -- TEST PROGRAM: the reference demo's CCD loop
// (CCDDemo/mainTorusknot.cpp:312-321) written against the source-compatible
// headers in include/c2a/ and linked with libc2a_amd.so.  Reads two triangle
// soups and a list of pose quadruples, runs C2A_Solve on each query in order
// with the demo's last_tri protocol, and writes the outputs for comparison.
//
// input : int32 n1, n1*9 doubles, int32 n2, n2*9 doubles, int32 nq,
//         nq*4*12 doubles (start/end pose of object 1, then of object 2;
//         row-major R then T)
// output: per query 2 doubles (toc, distance), 5 int32 (iterations,
//         contacts, collision free, last_tri of each model), 24 doubles
//         (contact poses of both objects, R then T)
#include <stdio.h>
#include <stdlib.h>

#include <cmath>
#include <vector>

#include "c2a/C2A.h"

static bool read_model(FILE* f, C2A_Model& m) {
  int n = 0;
  if (fread(&n, 4, 1, f) != 1 || n <= 0) return false;
  std::vector<double> t((size_t)n * 9);
  if (fread(t.data(), 8, t.size(), f) != t.size()) return false;
  m.BeginModel();
  for (int i = 0; i < n; i++) m.AddTri(&t[9 * i], &t[9 * i + 3], &t[9 * i + 6], i);
  return m.EndModel() == PQP_OK;
}

static void to_transform(const double* p, Transform& t) {
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) t.Rotation()[i][j] = p[3 * i + j];
    t.Translation()[i] = p[9 + i];
  }
}

static void put_transform(const Transform& t, double* o) {
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) o[3 * i + j] = t.Rotation()[i][j];
    o[9 + i] = t.Translation()[i];
  }
}

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    fprintf(stderr, "usage: solve_demo input.bin output.bin [reuse]\n");
    return 2;
  }
  // reuse: dres (and its last_triA/B) carries over from call to call instead
  // of being seeded from the models' last_tri before each call
  const bool reuse = argc == 4;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  C2A_Model obj1, obj2;
  if (!read_model(f, obj1) || !read_model(f, obj2)) return 3;
  int nq = 0;
  if (fread(&nq, 4, 1, f) != 1) return 3;
  std::vector<double> poses((size_t)nq * 48);
  if (fread(poses.data(), 8, poses.size(), f) != poses.size()) return 3;
  fclose(f);
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  C2A_TimeOfContactResult dres;
  for (int q = 0; q < nq; q++) {
    Transform trans00, trans01, trans10, trans11, trans0, trans1;
    to_transform(&poses[(size_t)q * 48], trans00);
    to_transform(&poses[(size_t)q * 48 + 12], trans01);
    to_transform(&poses[(size_t)q * 48 + 24], trans10);
    to_transform(&poses[(size_t)q * 48 + 36], trans11);
    trans0.Identity();
    trans1.Identity();
    PQP_REAL toc = 0;
    int nItr = 0, nContact = 0;
    if (!reuse || q == 0) {
      dres.last_triA = obj1.last_tri;
      dres.last_triB = obj2.last_tri;
    }
    C2A_Result r = C2A_Solve(&trans00, &trans01, &obj1, &trans10, &trans11, &obj2, trans0, trans1,
                             toc, nItr, nContact, 0.0, dres);
    if (r != TOCFound) return 4;
    if (!dres.collisionfree) {
      // the contact query at the TOC pose (C2A.cpp:2433-2435, 1727-1775):
      // one ContactF per counted pair, each within 2 d + 0.001, mint 1,
      // zero error schedule, res->R/T at the TOC pose
      if ((int)dres.cont_l.size() != nContact) return 5;
      ContactFList sorted = dres.cont_l;
      sorted.sort();   // ContactF::operator<
      double prev = -1.0;
      for (ContactFListIterator it = sorted.begin(); it != sorted.end(); ++it) {
        if (it->Distance < prev || it->Distance > 2 * dres.distance + 0.001) return 6;
        if (it->FeatureType_A < 1 || it->FeatureType_A > 3) return 6;
        prev = it->Distance;
      }
      if (dres.mint != 1.0 || dres.abs_err != 0.0 || dres.rel_err != 0.0) return 7;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double v = 0;
          for (int k = 0; k < 3; k++) v += trans0.Rotation()[k][i] * trans1.Rotation()[k][j];
          if (std::fabs(v - dres.R[i][j]) > 1e-12) return 8;
        }
    }
    double d2[2] = {toc, dres.distance};
    int i5[7] = {nItr, nContact, dres.collisionfree ? 1 : 0, obj1.tri_index(obj1.last_tri),
                 obj2.tri_index(obj2.last_tri), obj1.tri_index(dres.last_triA),
                 obj2.tri_index(dres.last_triB)};
    double tr[24];
    put_transform(trans0, tr);
    put_transform(trans1, tr + 12);
    fwrite(d2, 8, 2, o);
    fwrite(i5, 4, 7, o);
    fwrite(tr, 8, 24, o);
  }
  fclose(o);
  printf("solve_demo: %d queries\n", nq);
  return 0;
}
