// ccd_types.h -- device-side data of the CCD kernel (plain structs shared by
// the kernel, the host launcher and the host-side test harness).
#pragma once

#include <stdint.h>

#include "../../include/c2a_amd.h"
#include "bvh.h"
#include "c2a_math.h"

namespace c2a {


enum : int { MODE_QUERY = 0, MODE_STEP = 1, MODE_CONTACT = 2 };

struct MeshDev {
  const GpuNode* nodes;
  const double* tris;      // build order, 9 doubles each
  int32_t ntri;
  int32_t depth;
};

// One query as the kernel reads it: both motions precomputed on the host.
struct QueryDev {
  Motion mo[2];
  int32_t mesh[2];
  int32_t seed[2];
};

// One explicit CA step (backs C2A_TimeOfContactStep).
struct StepDev {
  Motion mo[2];
  double R1[9], T1[3], R2[9], T2[3];
  double mint_prev, UpboundTOC;
  int32_t mesh[2];
  int32_t seed[2];
  int32_t last[2];
  int32_t numCA;
  int32_t pad;
};

// One contact enumeration at explicit poses (backs C2A_QueryContact).
struct ContactDev {
  double R1[9], T1[3], R2[9], T2[3];
  double threshold;
  int32_t mesh[2];
};

struct KArgs {
  const MeshDev* meshes;
  const QueryDev* q;
  const StepDev* step_in;
  c2a_result* out;
  c2a_step_result* step_out;
  const ContactDev* contact_in;
  int32_t* counts_out;
  int64_t* counter;        // work-queue ticket, zeroed before each launch
  int64_t nq;
  double tol_d, tol_t;
  double contact_scale, contact_offset;
  int32_t max_iter;
  int32_t with_contacts;
  int32_t mode;
  int32_t pad;
};

}  // namespace c2a
