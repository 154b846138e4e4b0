// Instantiation of the step kernels for one (lattice, collision) pair.  Each
// step_<lattice>_<collision>.hip expands these macros once, so the library's
// translation units compile in parallel; launch.h declares the functions.
#pragma once

#include "launch.h"
#include "step_impl.h"

#define AERO_BATCH_CASE(L, COLL, ST, KMODE)                                                            \
  case ST:                                                                                             \
    return thermal ? launch_batch<L, COLL, ST, true, KMODE>(cases, tiles, ntiles, k, s, e0, e1)         \
                   : launch_batch<L, COLL, ST, false, KMODE>(cases, tiles, ntiles, k, s, e0, e1);

// Batched fused step (every case of a group, one launch per step); KMODE 0
// for the plain collision, 2 for the local Smagorinsky rates.
#define AERO_INST_BATCH_K(L, COLL, FN, KMODE)                                                  \
  hipError_t FN(int st, bool thermal, const BatchCase* cases, const int2* tiles, int ntiles,  \
                int k, hipStream_t s, void* e0, void* e1) {                                   \
    switch (st) {                                                                              \
      AERO_BATCH_CASE(L, COLL, ST_F64, KMODE)                                                  \
      AERO_BATCH_CASE(L, COLL, ST_F32, KMODE)                                                  \
      AERO_BATCH_CASE(L, COLL, ST_F16S, KMODE)                                                 \
      AERO_BATCH_CASE(L, COLL, ST_BF16S, KMODE)                                                \
    }                                                                                          \
    return hipErrorInvalidValue;                                                               \
  }
#define AERO_INST_BATCH(L, COLL, NAME) AERO_INST_BATCH_K(L, COLL, launch_batch_##NAME, 0)
#define AERO_INST_BATCH_LES(L, COLL, NAME) AERO_INST_BATCH_K(L, COLL, launch_batch_les_##NAME, 2)

#define AERO_SPLIT_CASE(L, COLL, ST) \
  case ST:                           \
    return thermal ? launch_step<L, COLL, ST, true, 1>(p, s) : launch_step<L, COLL, ST, false, 1>(p, s);

// Split step for the finite-difference LES pass (BGK and TRT only: the
// reference's MRT ignores the LES relaxation field).
#define AERO_INST_SPLIT(L, COLL, NAME)                                                         \
  hipError_t launch_split_##NAME(int st, bool thermal, const StepParams& p, hipStream_t s) {  \
    switch (st) {                                                                              \
      AERO_SPLIT_CASE(L, COLL, ST_F64)                                                         \
      AERO_SPLIT_CASE(L, COLL, ST_F32)                                                         \
      AERO_SPLIT_CASE(L, COLL, ST_F16S)                                                        \
      AERO_SPLIT_CASE(L, COLL, ST_BF16S)                                                       \
    }                                                                                          \
    return hipErrorInvalidValue;                                                               \
  }

