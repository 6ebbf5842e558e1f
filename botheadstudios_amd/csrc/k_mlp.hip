// Learned smoothing-length seed (placeholder until the network kernels land): the seed is the previous h.
#include "mlp.h"

int mlp_alloc(MlpDev& m, int64_t cap) {
  m.cap = cap;
  return 0;
}
void mlp_free(MlpDev& m) { m.cap = 0; }
void mlp_seed_h(hipStream_t, MlpDev&, int64_t, const MasterBufs&, const float*) {}
void mlp_record(hipStream_t, MlpDev&, int64_t, const MasterBufs&) {}
