// static_demo.cpp -- TEST PROGRAM: static queries through the
// source-compatible interface (C2A.h:241-267 and C2A_Tolerance,
// C2A_PQP.cpp:1762) linked with libc2a_amd.so.  One kind of query per run;
// the models' last_tri carries from one call to the next, as in the
// reference.
//
// input : int32 n1, n1*9 doubles, int32 n2, n2*9 doubles, int32 kind,
//         int32 qsize_or_flag, double param, int32 nq, nq*24 doubles (pose
//         of model 1, then of model 2; row-major R then T)
// kind 0  C2A_Distance(rel = abs = 0, qsize)        -> 7 doubles (distance,
//         p1, p2), 6 int32 (bv tests, tri tests, t1, t2, last_tri of each)
// kind 1  C2A_Distance_At_ContactSpace(param)      -> the same record
// kind 2  C2A_Tolerance(param, qsize)               -> 3 int32 (closer, bv
//         tests, tri tests), 7 doubles (distance, p1, p2)
// kind 3  C2A_Collide(PQP_CollideResult, flag)      -> 3 int32 (pairs, bv
//         tests, tri tests), then 2 int32 (id1, id2) per pair
#include <stdio.h>
#include <string.h>

#include <vector>

#include "c2a/C2A.h"

static bool read_model(FILE* f, C2A_Model& m) {
  int n = 0;
  if (fread(&n, 4, 1, f) != 1 || n <= 0) return false;
  std::vector<double> t((size_t)n * 9);
  if (fread(t.data(), 8, t.size(), f) != t.size()) return false;
  m.BeginModel();
  for (int i = 0; i < n; i++) m.AddTri(&t[9 * i], &t[9 * i + 3], &t[9 * i + 6], 1000 + i);
  return m.EndModel() == PQP_OK;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: static_demo input.bin output.bin\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  C2A_Model obj1, obj2;
  if (!read_model(f, obj1) || !read_model(f, obj2)) return 3;
  int kind = 0, qf = 0, nq = 0;
  double param = 0.0;
  if (fread(&kind, 4, 1, f) != 1 || fread(&qf, 4, 1, f) != 1 || fread(&param, 8, 1, f) != 1 ||
      fread(&nq, 4, 1, f) != 1)
    return 3;
  std::vector<double> poses((size_t)nq * 24);
  if (fread(poses.data(), 8, poses.size(), f) != poses.size()) return 3;
  fclose(f);
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  PQP_CollideResult coll;   // reused across calls, as a caller would
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
    if (kind == 0 || kind == 1) {
      C2A_DistanceResult res;
      memset(&res, 0, sizeof(res));
      const int rc = kind == 0 ? C2A_Distance(&res, R1, T1, &obj1, R2, T2, &obj2, 0.0, 0.0, qf)
                               : C2A_Distance_At_ContactSpace(&res, R1, T1, &obj1, R2, T2, &obj2,
                                                              param);
      if (rc != PQP_OK) return 4;
      double d7[7] = {res.Distance(), res.P1()[0], res.P1()[1], res.P1()[2],
                      res.P2()[0],    res.P2()[1], res.P2()[2]};
      int i6[6] = {res.NumBVTests(), res.NumTriTests(), res.t1, res.t2,
                   obj1.tri_index(obj1.last_tri), obj2.tri_index(obj2.last_tri)};
      fwrite(d7, 8, 7, o);
      fwrite(i6, 4, 6, o);
    } else if (kind == 2) {
      PQP_ToleranceResult res;
      memset(&res, 0, sizeof(res));
      if (C2A_Tolerance(&res, R1, T1, &obj1, R2, T2, &obj2, param, qf) != PQP_OK) return 4;
      int i3[3] = {res.CloserThanTolerance(), res.NumBVTests(), res.NumTriTests()};
      double d7[7] = {res.Distance(), res.P1()[0], res.P1()[1], res.P1()[2],
                      res.P2()[0],    res.P2()[1], res.P2()[2]};
      fwrite(i3, 4, 3, o);
      fwrite(d7, 8, 7, o);
    } else {
      if (C2A_Collide(&coll, R1, T1, &obj1, R2, T2, &obj2, qf) != PQP_OK) return 4;
      int i3[3] = {coll.NumPairs(), coll.NumBVTests(), coll.NumTriTests()};
      fwrite(i3, 4, 3, o);
      for (int k = 0; k < coll.NumPairs(); k++) {
        int ids[2] = {coll.Id1(k), coll.Id2(k)};
        fwrite(ids, 4, 2, o);
      }
      if (coll.Colliding() != (coll.NumPairs() > 0)) return 5;
    }
  }
  fclose(o);
  return 0;
}
