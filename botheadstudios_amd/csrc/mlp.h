// Learned smoothing-length seed: a 6->32->32->1 GELU network with BF16 weights and FP32 accumulation.
//
// Inputs (from the previous force evaluation's state, before the Newton solve of a post-drift evaluation):
//   x0 = 100 div v dt (percent volume change over the drift), x1 = ln(rho / rho0_mat), x2 = (n_ngb - 47) / 47,
//   x3 = |grad rho| * h / rho, x4 = ln(h / h_ref_mat) with h_ref = eta (m / rho0)^(1/3), x5 = material id.
// Output y: the correction to the continuity-equation estimate, i.e. the seed is
//   h_prev * exp((MLP_PHYS * x0 + y) / MLP_YSCALE)
// (h ~ rho^(-1/3) and d ln rho = -div v dt give ln(h_new / h_prev) ~ div v dt / 3 to first order), clamped to
// [h_prev / MLP_CLAMP, h_prev * MLP_CLAMP]. The Newton solve accepts or corrects the seed.
// Zero weights give the continuity estimate alone.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"

constexpr int MLP_IN = 6;
constexpr int MLP_H = 32;
// Parameter layout (FP32 master / BF16 copy): W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3.
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_H * MLP_IN;  // 192
constexpr int MLP_W2 = MLP_B1 + MLP_H;           // 224
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;   // 1248
constexpr int MLP_W3 = MLP_B2 + MLP_H;           // 1280
constexpr int MLP_B3 = MLP_W3 + MLP_H;           // 1312
constexpr int MLP_NPARAM = MLP_B3 + 1;           // 1313
constexpr float MLP_YSCALE = 100.0f;
constexpr float MLP_PHYS = 1.0f / 3.0f;
constexpr float MLP_CLAMP = 1.2f;
constexpr int MLP_SAMPLE = 8;  // floats per recorded sample: 6 features, target, spare

struct MlpDev {
  bool record = false;   // store (features, converged h) samples at each post-drift force evaluation
  bool trained = false;  // weights present; otherwise the seed is the previous h
  int64_t cap = 0;
  float* w = nullptr;       // FP32 master weights
  uint16_t* wb = nullptr;   // BF16 weights used by inference and training forward passes
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* gpart = nullptr;   // per-block gradient partials
  float* grad = nullptr;    // reduced gradient + loss
  float* xbuf = nullptr;    // features of the current evaluation (n * 6), for recording
  float* samples = nullptr; // recorded samples (sample_cap * MLP_SAMPLE)
  int64_t sample_cap = 0;
  int64_t nsamples = 0;
  int64_t adam_t = 0;
  int gblocks = 0;
};

int mlp_alloc(MlpDev& m, int64_t cap);
void mlp_free(MlpDev& m);
// Writes the Newton initial guess into D.h and the searched radius hs = min(seed * margin, max(hcap, seed)),
// with each particle's segment drift dt and cap.
void mlp_seed_h(hipStream_t s, MlpDev& m, int64_t n, const MasterBufs& D, float* hs, const void* eos,
                const SegView& sv, float eta, float margin);
// Recording: features before the solve, then the converged target after it.
void mlp_record_features(hipStream_t s, MlpDev& m, int64_t n, const MasterBufs& D, const void* eos, const SegView& sv,
                         float eta);
void mlp_record_targets(hipStream_t s, MlpDev& m, int64_t n, const MasterBufs& D);
// Gradient of the mean squared error over n samples into m.grad[0..NPARAM), mean loss in grad[NPARAM].
// Sample k is row (off + k * step) % rows of x (row stride xstride floats) and y (stride ystride).
void mlp_gradient(hipStream_t s, MlpDev& m, const float* x, int64_t xstride, const float* y, int64_t ystride,
                  int64_t n, int64_t rows, int64_t off, int64_t step);
void mlp_adam(hipStream_t s, MlpDev& m, float lr);
void mlp_forward(hipStream_t s, MlpDev& m, const float* x, int64_t n, float* out);
void mlp_sync_bf16(hipStream_t s, MlpDev& m);
