// Learned smoothing-length seed: a 6->32->32->1 GELU network with BF16 weights and FP32 accumulation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"

struct MlpDev {
  bool record = false;   // store (features, converged h) samples at each post-drift force evaluation
  bool trained = false;  // weights uploaded or trained; otherwise the seed is the previous h
  int64_t cap = 0;
};

int mlp_alloc(MlpDev& m, int64_t cap);
void mlp_free(MlpDev& m);
// Writes the Newton initial guess into D.h (clamped to the searched radius hs).
void mlp_seed_h(hipStream_t s, MlpDev& m, int64_t n, const MasterBufs& D, const float* hs);
void mlp_record(hipStream_t s, MlpDev& m, int64_t n, const MasterBufs& D);
