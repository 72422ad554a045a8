// distance_demo.cpp -- TEST PROGRAM: a sequence of static distance queries
// through the source-compatible C2A_Distance (C2A.h:255-259) linked with
// libc2a_amd.so.  The models' last_tri carries from one call to the next, as
// in the reference, where it seeds the next call's first bound.
//
// input : int32 n1, n1*9 doubles, int32 n2, n2*9 doubles, double rel_err,
//         double abs_err, int32 nq, nq*24 doubles (pose of model 1, then of
//         model 2; row-major R then T)
// output: per query 7 doubles (distance, p1, p2), 6 int32 (bv tests, tri
//         tests, t1, t2, last_tri index of each model), 12 doubles (R, T)
#include <stdio.h>

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

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: distance_demo input.bin output.bin\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  C2A_Model obj1, obj2;
  if (!read_model(f, obj1) || !read_model(f, obj2)) return 3;
  double err[2];
  int nq = 0;
  if (fread(err, 8, 2, f) != 2 || fread(&nq, 4, 1, f) != 1) return 3;
  std::vector<double> poses((size_t)nq * 24);
  if (fread(poses.data(), 8, poses.size(), f) != poses.size()) return 3;
  fclose(f);
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  for (int q = 0; q < nq; q++) {
    const double* p = &poses[(size_t)q * 24];
    PQP_REAL R1[3][3], T1[3], R2[3][3], T2[3];
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) {
        R1[i][j] = p[3 * i + j];
        R2[i][j] = p[12 + 3 * i + j];
      }
      T1[i] = p[9 + i];
      T2[i] = p[21 + i];
    }
    C2A_DistanceResult res;
    if (C2A_Distance(&res, R1, T1, &obj1, R2, T2, &obj2, err[0], err[1]) != PQP_OK) return 4;
    double d7[7] = {res.Distance(), res.P1()[0], res.P1()[1], res.P1()[2],
                    res.P2()[0],    res.P2()[1], res.P2()[2]};
    int i6[6] = {res.NumBVTests(), res.NumTriTests(), res.t1, res.t2,
                 obj1.tri_index(obj1.last_tri), obj2.tri_index(obj2.last_tri)};
    double rt[12];
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) rt[3 * i + j] = res.R[i][j];
      rt[9 + i] = res.T[i];
    }
    fwrite(d7, 8, 7, o);
    fwrite(i6, 4, 6, o);
    fwrite(rt, 8, 12, o);
  }
  fclose(o);
  return 0;
}
