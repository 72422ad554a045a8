// C2A_Internal.h -- the model type of the reference's CCD interface
// (C2A/C2A_Internal.h:32-81 over PQP_Model): BeginModel / AddTri / EndModel,
// the triangles in hierarchy order and the last_tri pointer the CCD query
// reads and updates.  The hierarchy itself lives in the MI355X engine.
#ifndef C2A_COMPAT_INTERNAL_H
#define C2A_COMPAT_INTERNAL_H

#include <vector>

typedef double PQP_REAL;

// PQP return codes the model calls use.
enum { PQP_OK = 0, PQP_ERR_MODEL_OUT_OF_MEMORY = -1, PQP_ERR_UNPROCESSED_MODEL = -3,
       PQP_ERR_BUILD_OUT_OF_SEQUENCE = -4, PQP_ERR_BUILD_EMPTY_MODEL = -5 };

struct c2a_mesh;

// A triangle as PQP stores it (p1, p2, p3, caller id).
struct Tri {
  PQP_REAL p1[3];
  PQP_REAL p2[3];
  PQP_REAL p3[3];
  int id;
};

class C2A_Model {
 public:
  C2A_Model();
  ~C2A_Model();
  C2A_Model(const C2A_Model&) = delete;
  C2A_Model& operator=(const C2A_Model&) = delete;

  int BeginModel(int n = 8);
  int AddTri(const PQP_REAL* p1, const PQP_REAL* p2, const PQP_REAL* p3, int id);
  int AddTri(const PQP_REAL* p1, const PQP_REAL* p2, const PQP_REAL* p3, int i, int i1, int i2,
             int i3);
  int EndModel();

  Tri* tris;           // hierarchy (build) order after EndModel
  int num_tris;
  Tri* last_tri;       // closest triangle of the last CCD step (tris after EndModel)
  int build_state;     // 0 empty, 1 building, 2 processed

  c2a_mesh* engine_mesh() const { return mesh_; }
  int tri_index(const Tri* t) const { return t ? (int)(t - tris) : 0; }

 private:
  std::vector<PQP_REAL> xyz_;
  std::vector<int> ids_, vidx_;
  std::vector<Tri> built_;
  c2a_mesh* mesh_;
};

#endif
