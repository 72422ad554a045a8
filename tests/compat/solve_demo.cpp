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
  if (argc != 3) {
    fprintf(stderr, "usage: solve_demo input.bin output.bin\n");
    return 2;
  }
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
    dres.last_triA = obj1.last_tri;
    dres.last_triB = obj2.last_tri;
    C2A_Result r = C2A_Solve(&trans00, &trans01, &obj1, &trans10, &trans11, &obj2, trans0, trans1,
                             toc, nItr, nContact, 0.0, dres);
    if (r != TOCFound) return 4;
    double d2[2] = {toc, dres.distance};
    int i5[5] = {nItr, nContact, dres.collisionfree ? 1 : 0, obj1.tri_index(obj1.last_tri),
                 obj2.tri_index(obj2.last_tri)};
    double tr[24];
    put_transform(trans0, tr);
    put_transform(trans1, tr + 12);
    fwrite(d2, 8, 2, o);
    fwrite(i5, 4, 5, o);
    fwrite(tr, 8, 24, o);
  }
  fclose(o);
  printf("solve_demo: %d queries\n", nq);
  return 0;
}
