// CPU oracle for the SPH giant-impact step — TEST INFRASTRUCTURE ONLY (checker, never the product).
// Restates the reference's CPU twins and LBVH arithmetic; each function cites the reference file:line
// it follows (paths relative to the reference repository; ie/ = integrity-engine/).
// Built with -ffp-contract=off so every FP32 expression rounds exactly as written.
#include "sph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kG = 6.674e-11;   // ie/crates/engine/src/orbit.rs:18
constexpr float kG32 = 6.674e-11f; // ie/shaders/bh_gravity.wgsl:11
constexpr double kPi = 3.14159265358979323846;

inline void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

struct Till {
  double rho0, a, b, A, B, E0, Eiv, Ecv, alpha, beta;
  explicit Till(const double* e)
      : rho0(e[0]), a(e[1]), b(e[2]), A(e[3]), B(e[4]), E0(e[5]), Eiv(e[6]), Ecv(e[7]), alpha(e[8]),
        beta(e[9]) {}
};

// ie/crates/engine/src/eos.rs:52-83 (pressure, p_compressed, p_expanded), f64.
double pressure64(const Till& t, double rho, double u) {
  rho = std::max(rho, 1.0e-9);
  const double eta = rho / t.rho0;
  const double mu = eta - 1.0;
  const double omega_c = u / (t.E0 * eta * eta) + 1.0;
  const double p_c = (t.a + t.b / omega_c) * rho * u + t.A * mu + t.B * mu * mu;
  if (eta >= 1.0 || u <= t.Eiv) return p_c;
  const double omega_e = u / (t.E0 * eta * eta) + 1.0;
  const double z = t.rho0 / rho - 1.0;
  const double p_e = t.a * rho * u + (t.b * rho * u / omega_e + t.A * mu * std::exp(-t.beta * z)) *
                                         std::exp(-t.alpha * z * z);
  if (u >= t.Ecv) return p_e;
  return ((u - t.Eiv) * p_e + (t.Ecv - u) * p_c) / (t.Ecv - t.Eiv);
}

// ie/crates/engine/src/eos.rs:88-96 — central differences with 1e-4 steps, clamped >= 0.
double sound_speed_sq64(const Till& t, double rho, double u) {
  rho = std::max(rho, 1.0e-9);
  const double dr = rho * 1.0e-4;
  const double du = std::fabs(u) * 1.0e-4 + 1.0;
  const double dp_drho = (pressure64(t, rho + dr, u) - pressure64(t, rho - dr, u)) / (2.0 * dr);
  const double dp_du = (pressure64(t, rho, u + du) - pressure64(t, rho, u - du)) / (2.0 * du);
  const double p = pressure64(t, rho, u);
  return std::max(dp_drho + p / (rho * rho) * dp_du, 0.0);
}

// ie/tools/sph-verify/src/main.rs:93-100 (the WGSL finite-difference steps, sph_step.wgsl:90-100), f64.
double sound_speed_wgsl64(const Till& t, double rho, double u) {
  const double r = std::max(rho, 1.0e-9);
  const double dr = r * 1.0e-3;
  const double dp = (pressure64(t, r + dr, u) - pressure64(t, r - dr, u)) / (2.0 * dr);
  const double du = std::fabs(u) * 1.0e-3 + 1.0;
  const double dpu = (pressure64(t, r, u + du) - pressure64(t, r, u - du)) / (2.0 * du);
  const double pp = pressure64(t, r, u);
  return std::sqrt(std::max(dp + pp / (r * r) * dpu, 0.0));
}

// ie/crates/engine/src/atmosphere.rs:39-49.
double sph_w64(double r, double h) {
  const double q = r / h;
  const double sigma = 8.0 / (kPi * h * h * h);
  if (q < 0.5) return sigma * (6.0 * (q * q * q - q * q) + 1.0);
  if (q < 1.0) {
    const double t = 1.0 - q;
    return sigma * 2.0 * t * t * t;
  }
  return 0.0;
}

// ie/crates/engine/src/atmosphere.rs:52-62.
double sph_dw64(double r, double h) {
  const double q = r / h;
  const double sigma = 8.0 / (kPi * h * h * h * h);
  if (q < 0.5) return sigma * (18.0 * q * q - 12.0 * q);
  if (q < 1.0) return sigma * -6.0 * (1.0 - q) * (1.0 - q);
  return 0.0;
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

// ---------------------------------------------------------------------------------------------
// neighbors.rs:28-85 — uniform hash grid; pairs (i<j) in the 27 surrounding cells, each once.
struct CellKey {
  int32_t x, y, z;
  bool operator==(const CellKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct CellHash {
  size_t operator()(const CellKey& k) const {
    uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// neighbors.rs:35-37: floor(p / cell) as i32.
inline CellKey key_of(const double* p, double cell) {
  return CellKey{(int32_t)std::floor(p[0] / cell), (int32_t)std::floor(p[1] / cell),
                 (int32_t)std::floor(p[2] / cell)};
}

struct Grid {
  double cell = 1.0;
  int64_t n = 0;
  bool brute = true;
  std::unordered_map<CellKey, std::vector<int64_t>, CellHash> cells;
  // neighbors.rs:41-54 (BRUTE_BELOW = 512, neighbors.rs:23).
  Grid(const double* xyz, int64_t n_, double cell_, bool force_brute) : n(n_) {
    cell = std::max(cell_, 1.0e-9);
    brute = force_brute || n < 512;
    if (!brute)
      for (int64_t i = 0; i < n; ++i) cells[key_of(xyz + 3 * i, cell)].push_back(i);
  }
  // neighbors.rs:60-85.
  template <class F>
  void for_each_pair(const double* xyz, F&& f) const {
    if (brute) {
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) f(i, j);
      return;
    }
    for (int64_t i = 0; i < n; ++i) {
      const CellKey c = key_of(xyz + 3 * i, cell);
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            auto it = cells.find(CellKey{c.x + dx, c.y + dy, c.z + dz});
            if (it == cells.end()) continue;
            for (int64_t j : it->second)
              if (j > i) f(i, j);
          }
    }
  }
};

// ---------------------------------------------------------------------------------------------
// bhtree.rs:25-190 — f64 octree Barnes-Hut with MAC (2*half)/dist < theta.
struct Octree {
  struct Node {
    double cx, cy, cz, half;
    double comx, comy, comz, mass;
    int64_t child[8];
    std::vector<int64_t> leaf;
  };
  std::vector<Node> nodes;
  double theta, soft2;
  int64_t n;
  int leaf_cap;

  int64_t build_node(const std::vector<int64_t>& idx, double cx, double cy, double cz, double half,
                     const double* xyz, const double* m, int depth) {
    const int64_t id = (int64_t)nodes.size();
    double ms = 0, sx = 0, sy = 0, sz = 0;
    for (int64_t i : idx) {
      ms += m[i];
      sx += xyz[3 * i] * m[i];
      sy += xyz[3 * i + 1] * m[i];
      sz += xyz[3 * i + 2] * m[i];
    }
    Node nd;
    nd.cx = cx; nd.cy = cy; nd.cz = cz; nd.half = half;
    if (ms > 0) { nd.comx = sx / ms; nd.comy = sy / ms; nd.comz = sz / ms; }
    else { nd.comx = cx; nd.comy = cy; nd.comz = cz; }
    nd.mass = ms;
    for (auto& c : nd.child) c = -1;
    nodes.push_back(nd);
    // bhtree.rs:85 (leaf_cap = 1 is the reference's idx.len() <= 1; MAX_DEPTH = 28, bhtree.rs:23).
    if ((int64_t)idx.size() <= leaf_cap || depth >= 28) {
      nodes[id].leaf = idx;
      return id;
    }
    std::vector<int64_t> oct[8];
    for (int64_t i : idx) {
      const int o = (xyz[3 * i] >= cx ? 1 : 0) | (xyz[3 * i + 1] >= cy ? 2 : 0) | (xyz[3 * i + 2] >= cz ? 4 : 0);
      oct[o].push_back(i);
    }
    const double ch = half * 0.5;
    for (int o = 0; o < 8; ++o) {
      if (oct[o].empty()) continue;
      const int64_t cid = build_node(oct[o], cx + ((o & 1) ? ch : -ch), cy + ((o & 2) ? ch : -ch),
                                     cz + ((o & 4) ? ch : -ch), ch, xyz, m, depth + 1);
      nodes[id].child[o] = cid;
    }
    return id;
  }

  // bhtree.rs:45-62.
  Octree(const double* xyz, const double* m, int64_t n_, double theta_, double soft, int leaf_cap_)
      : theta(theta_), soft2(soft * soft), n(n_), leaf_cap(leaf_cap_ < 1 ? 1 : leaf_cap_) {
    if (n < 1024) return;  // BRUTE_BELOW, bhtree.rs:21
    double lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (int64_t i = 1; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], xyz[3 * i + k]);
        hi[k] = std::max(hi[k], xyz[3 * i + k]);
      }
    const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    const double half = std::max(ext * 0.5, 1.0e-9) * 1.0001;
    std::vector<int64_t> all(n);
    std::iota(all.begin(), all.end(), 0);
    nodes.reserve(2 * n);
    build_node(all, (lo[0] + hi[0]) * 0.5, (lo[1] + hi[1]) * 0.5, (lo[2] + hi[2]) * 0.5, half, xyz, m, 0);
  }

  // bhtree.rs:164-190.
  void accel_on(int64_t i, int64_t node, const double* xyz, const double* m, double* a) const {
    const Node& nd = nodes[node];
    if (!nd.leaf.empty()) {
      for (int64_t j : nd.leaf) {
        if (j == i) continue;
        const double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1],
                     dz = xyz[3 * j + 2] - xyz[3 * i + 2];
        const double r2 = dx * dx + dy * dy + dz * dz + soft2;
        const double g = kG * m[j] / (r2 * std::sqrt(r2));
        a[0] += dx * g; a[1] += dy * g; a[2] += dz * g;
      }
      return;
    }
    const double dx = nd.comx - xyz[3 * i], dy = nd.comy - xyz[3 * i + 1], dz = nd.comz - xyz[3 * i + 2];
    const double r2 = dx * dx + dy * dy + dz * dz + soft2;
    const double dist = std::sqrt(r2);
    if ((2.0 * nd.half) / dist < theta) {
      const double g = kG * nd.mass / (r2 * dist);
      a[0] += dx * g; a[1] += dy * g; a[2] += dz * g;
      return;
    }
    for (int c = 0; c < 8; ++c)
      if (nd.child[c] >= 0) accel_on(i, nd.child[c], xyz, m, a);
  }

  // bhtree.rs:117-134 (brute-force below BRUTE_BELOW).
  void accelerations(const double* xyz, const double* m, double* acc, int nthreads) const {
    set_threads(nthreads);
    if (nodes.empty()) {
#pragma omp parallel for schedule(dynamic, 64)
      for (int64_t i = 0; i < n; ++i) {
        double a[3] = {0, 0, 0};
        for (int64_t j = 0; j < n; ++j) {
          if (i == j) continue;
          const double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1],
                       dz = xyz[3 * j + 2] - xyz[3 * i + 2];
          const double r2 = dx * dx + dy * dy + dz * dz + soft2;
          const double g = kG * m[j] / (r2 * std::sqrt(r2));
          a[0] += dx * g; a[1] += dy * g; a[2] += dz * g;
        }
        acc[3 * i] = a[0]; acc[3 * i + 1] = a[1]; acc[3 * i + 2] = a[2];
      }
      return;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
      double a[3] = {0, 0, 0};
      accel_on(i, 0, xyz, m, a);
      acc[3 * i] = a[0]; acc[3 * i + 1] = a[1]; acc[3 * i + 2] = a[2];
    }
  }
};

// Candidate neighbour lists within radius rmax (exact, from a grid of cell rmax). Used by the
// Newton h-solve, which needs r < h for a varying h <= rmax.
void candidates(const double* xyz, int64_t n, double rmax, std::vector<std::vector<std::pair<double, int64_t>>>& out) {
  out.assign(n, {});
  Grid g(xyz, n, rmax, false);
  const double r2max = rmax * rmax;
  g.for_each_pair(xyz, [&](int64_t i, int64_t j) {
    const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                 dz = xyz[3 * i + 2] - xyz[3 * j + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < r2max) {
      const double r = std::sqrt(r2);
      out[i].push_back({r, j});
      out[j].push_back({r, i});
    }
  });
  for (auto& v : out) std::sort(v.begin(), v.end());
}

// rho~(h) = sum_j m_j W(r_ij, h) (self included) and its h-derivative; dW/dh = -(3W + r dW/dr)/h.
inline void own_density(const std::vector<std::pair<double, int64_t>>& cand, const double* m, int64_t i, double h,
                        double& rho, double& drho) {
  rho = m[i] * sph_w64(0.0, h);
  drho = -3.0 * rho / h;
  for (const auto& c : cand) {
    if (c.first >= h) break;
    const double w = sph_w64(c.first, h);
    const double dw = sph_dw64(c.first, h);
    rho += m[c.second] * w;
    drho += m[c.second] * (-(3.0 * w + c.first * dw) / h);
  }
}

}  // namespace

extern "C" {

double or_pressure(const double* e, double rho, double u) { return pressure64(Till(e), rho, u); }
double or_sound_speed_sq(const double* e, double rho, double u) { return sound_speed_sq64(Till(e), rho, u); }
double or_sound_speed_wgsl(const double* e, double rho, double u) { return sound_speed_wgsl64(Till(e), rho, u); }
double or_sph_w(double r, double h) { return sph_w64(r, h); }
double or_sph_dw(double r, double h) { return sph_dw64(r, h); }

// ie/shaders/sph_step.wgsl:78-89 evaluated in FP32 (operator order as written there).
float or_pressure_f32(const float* e, float rho, float u) {
  const float rho0 = e[0], a = e[1], b = e[2], cap_a = e[3], cap_b = e[4], e0 = e[5], e_iv = e[6], e_cv = e[7],
              alpha = e[8], beta = e[9];
  const float r = std::max(rho, 1.0e-9f);
  const float eta = r / rho0;
  const float mu = eta - 1.0f;
  const float omega = u / (e0 * eta * eta) + 1.0f;
  const float p_c = (a + b / omega) * r * u + cap_a * mu + cap_b * mu * mu;
  if (eta >= 1.0f || u <= e_iv) return p_c;
  const float z = rho0 / r - 1.0f;
  const float p_e = a * r * u + (b * r * u / omega + cap_a * mu * std::exp(-beta * z)) * std::exp(-alpha * z * z);
  if (u >= e_cv) return p_e;
  return ((u - e_iv) * p_e + (e_cv - u) * p_c) / (e_cv - e_iv);
}

// ie/crates/engine/src/gpu_gravity.rs:495-502 / ie/shaders/bh_gravity.wgsl:82-89.
uint32_t or_expand_bits(uint32_t v) {
  v &= 0x000003ffu;
  v = (v * 0x00010001u) & 0xff0000ffu;
  v = (v * 0x00000101u) & 0x0f00f00fu;
  v = (v * 0x00000011u) & 0xc30c30c3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// ie/crates/engine/src/gpu_gravity.rs:336-343 (f32 min/max of every coordinate).
void or_bbox_f32(const float* xyz, int64_t n, float* lo, float* hi) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = std::numeric_limits<float>::infinity();
    hi[k] = -std::numeric_limits<float>::infinity();
  }
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], xyz[3 * i + k]);
      hi[k] = std::max(hi[k], xyz[3 * i + k]);
    }
}

// ie/crates/engine/src/gpu_gravity.rs:345-357 (bit-identical to bh_gravity.wgsl:91-98).
void or_morton_f32(const float* xyz, int64_t n, const float* lo, const float* hi, uint32_t* codes) {
  float ext[3];
  for (int k = 0; k < 3; ++k) ext[k] = std::max(hi[k] - lo[k], 1.0e-30f);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      float u = (xyz[3 * i + k] - lo[k]) / ext[k];
      u = std::min(std::max(u, 0.0f), 1.0f);
      float f = std::floor(u * 1024.0f);
      f = std::min(std::max(f, 0.0f), 1023.0f);
      q[k] = (uint32_t)f;
    }
    codes[i] = or_expand_bits(q[0]) * 4u + or_expand_bits(q[1]) * 2u + or_expand_bits(q[2]);
  }
}

// ie/crates/engine/src/gpu_gravity.rs:358-363: pairs.sort_unstable() on (code, index).
void or_sort_pairs(const uint32_t* codes, int64_t n, uint32_t* sorted_codes, uint32_t* order) {
  std::vector<uint64_t> key(n);
  for (int64_t i = 0; i < n; ++i) key[i] = ((uint64_t)codes[i] << 32) | (uint64_t)(uint32_t)i;
  std::sort(key.begin(), key.end());
  for (int64_t i = 0; i < n; ++i) {
    sorted_codes[i] = (uint32_t)(key[i] >> 32);
    order[i] = (uint32_t)(key[i] & 0xffffffffu);
  }
}

// ie/crates/engine/src/gpu_gravity.rs:364-366: cluster code = first code of each bucket of k.
void or_cluster_codes(const uint32_t* sorted_codes, int64_t n, int64_t k, uint32_t* ccodes) {
  const int64_t L = (n + k - 1) / k;
  for (int64_t c = 0; c < L; ++c) ccodes[c] = sorted_codes[c * k];
}

// ie/shaders/bh_gravity.wgsl:114-184 — Karras 2012 radix tree over L cluster codes. Internal nodes
// [0, L-1), leaves [L-1, 2L-1), root 0; parent of the root = 0xffffffff.
void or_karras(const uint32_t* cc, int64_t L, uint32_t* left, uint32_t* right, uint32_t* parent) {
  or_karras_ex(cc, L, 0, left, right, parent);
}

// fix_dups = 1: for a range of equal cluster codes split at the highest differing bit of the
// index-augmented key (the split consistent with delta()'s ranges) instead of the reference's
// (first + last) >> 1, which yields an invalid tree for runs of three or more equal codes.
void or_karras_ex(const uint32_t* cc, int64_t L, int32_t fix_dups, uint32_t* left, uint32_t* right,
                  uint32_t* parent) {
  const int64_t nn = 2 * L - 1;
  for (int64_t i = 0; i < nn; ++i) {
    left[i] = 0;
    right[i] = 0;
    parent[i] = 0xffffffffu;
  }
  auto delta = [&](int64_t i, int64_t j) -> int {
    if (j < 0 || j >= L) return -1;
    if (cc[i] == cc[j]) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(cc[i] ^ cc[j]);
  };
  for (int64_t i = 0; i < L - 1; ++i) {
    const int64_t d = delta(i, i + 1) > delta(i, i - 1) ? 1 : -1;
    const int dmin = delta(i, i - d);
    int64_t lmax = 2;
    while (delta(i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (delta(i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int64_t first = std::min(i, j), last = std::max(i, j);
    int64_t split = first;
    const uint32_t fc = cc[first], lc = cc[last];
    if (fc == lc) {
      if (fix_dups) {
        const int p = 31 - clz32((uint32_t)first ^ (uint32_t)last);
        split = (int64_t)((((uint32_t)first >> p) + 1u) << p) - 1;
      } else {
        split = (first + last) >> 1;
      }
    } else {
      const int cpfx = clz32(fc ^ lc);
      int64_t step = last - first;
      while (true) {
        step = (step + 1) >> 1;
        const int64_t ns = split + step;
        if (ns < last && clz32(fc ^ cc[ns]) > cpfx) split = ns;
        if (step <= 1) break;
      }
    }
    const uint32_t li = (split == first) ? (uint32_t)(L - 1 + split) : (uint32_t)split;
    const uint32_t ri = (split + 1 == last) ? (uint32_t)(L - 1 + split + 1) : (uint32_t)(split + 1);
    left[i] = li;
    right[i] = ri;
    parent[li] = (uint32_t)i;
    parent[ri] = (uint32_t)i;
  }
}

// ie/shaders/bh_gravity.wgsl:205-241 (leaf bucket sums, mass-weighted merge, AABB union), FP32,
// evaluated children-first.
void or_upsweep_f32(const float* sb, int64_t n, int64_t k, const uint32_t* left, const uint32_t* right, int64_t L,
                    float* com4, float* bmin3, float* bmax3) {
  const int64_t leaf0 = L - 1;
  for (int64_t c = 0; c < L; ++c) {
    const int64_t li = leaf0 + c;
    const int64_t start = c * k, end = std::min(start + k, n);
    float m = 0.f, mx = 0.f, my = 0.f, mz = 0.f;
    float lo[3] = {1.0e30f, 1.0e30f, 1.0e30f}, hi[3] = {-1.0e30f, -1.0e30f, -1.0e30f};
    for (int64_t s = start; s < end; ++s) {
      const float* b = sb + 4 * s;
      m += b[3];
      mx += b[0] * b[3];
      my += b[1] * b[3];
      mz += b[2] * b[3];
      for (int q = 0; q < 3; ++q) {
        lo[q] = std::min(lo[q], b[q]);
        hi[q] = std::max(hi[q], b[q]);
      }
    }
    com4[4 * li + 0] = m > 0.f ? mx / m : 0.f;
    com4[4 * li + 1] = m > 0.f ? my / m : 0.f;
    com4[4 * li + 2] = m > 0.f ? mz / m : 0.f;
    com4[4 * li + 3] = m;
    for (int q = 0; q < 3; ++q) {
      bmin3[3 * li + q] = lo[q];
      bmax3[3 * li + q] = hi[q];
    }
  }
  if (L < 2) return;
  // Post-order over internal nodes (explicit stack).
  std::vector<int64_t> order;
  order.reserve(L);
  std::vector<int64_t> st{0};
  while (!st.empty()) {
    const int64_t x = st.back();
    st.pop_back();
    order.push_back(x);
    if ((int64_t)left[x] < leaf0) st.push_back(left[x]);
    if ((int64_t)right[x] < leaf0) st.push_back(right[x]);
  }
  for (auto it = order.rbegin(); it != order.rend(); ++it) {
    const int64_t x = *it, l = left[x], r = right[x];
    const float* cl = com4 + 4 * l;
    const float* cr = com4 + 4 * r;
    const float m = cl[3] + cr[3];
    for (int q = 0; q < 3; ++q) com4[4 * x + q] = m > 0.f ? (cl[q] * cl[3] + cr[q] * cr[3]) / m : 0.f;
    com4[4 * x + 3] = m;
    for (int q = 0; q < 3; ++q) {
      bmin3[3 * x + q] = std::min(bmin3[3 * l + q], bmin3[3 * r + q]);
      bmax3[3 * x + q] = std::max(bmax3[3 * l + q], bmax3[3 * r + q]);
    }
  }
}

// ie/shaders/bh_gravity.wgsl:337-394 — one sorted particle's stack walk with the robust MAC
// dist >= |bmax-bmin|/theta + |centre-com| and exact leaf sums (s != t). cnt (may be null) receives the accepted
// nodes, the leaf-particle interactions and the opened nodes of the walk.
namespace {
void bh_walk_one(const float* sb, int64_t t, int64_t n, int64_t k, const uint32_t* left, const uint32_t* right,
                 const float* com4, const float* bmin3, const float* bmax3, int64_t L, float theta, float s2,
                 float* a3, int64_t* cnt) {
  const int64_t leaf0 = L - 1;
  const float px = sb[4 * t], py = sb[4 * t + 1], pz = sb[4 * t + 2];
  float ax = 0.f, ay = 0.f, az = 0.f;
  int64_t n_node = 0, n_pp = 0, n_open = 0;
  uint32_t stack[64];
  uint32_t sp = 0;
  stack[sp++] = 0;
  while (sp != 0) {
    const uint32_t ni = stack[--sp];
    if ((int64_t)ni >= leaf0) {
      const int64_t c = ni - leaf0;
      const int64_t start = c * k, end = std::min(start + k, n);
      for (int64_t s = start; s < end; ++s) {
        if (s == t) continue;
        const float dx = sb[4 * s] - px, dy = sb[4 * s + 1] - py, dz = sb[4 * s + 2] - pz;
        const float r2 = (dx * dx + dy * dy + dz * dz) + s2;
        const float g = kG32 * sb[4 * s + 3] / (r2 * std::sqrt(r2));
        ax += dx * g; ay += dy * g; az += dz * g;
        ++n_pp;
      }
      continue;
    }
    const float* c4 = com4 + 4 * ni;
    const float dx = c4[0] - px, dy = c4[1] - py, dz = c4[2] - pz;
    const float r2 = (dx * dx + dy * dy + dz * dz) + s2;
    const float dist = std::sqrt(r2);
    const float sx = bmax3[3 * ni] - bmin3[3 * ni], sy = bmax3[3 * ni + 1] - bmin3[3 * ni + 1],
                sz = bmax3[3 * ni + 2] - bmin3[3 * ni + 2];
    const float size = std::sqrt(sx * sx + sy * sy + sz * sz);
    const float ex = 0.5f * (bmin3[3 * ni] + bmax3[3 * ni]) - c4[0];
    const float ey = 0.5f * (bmin3[3 * ni + 1] + bmax3[3 * ni + 1]) - c4[1];
    const float ez = 0.5f * (bmin3[3 * ni + 2] + bmax3[3 * ni + 2]) - c4[2];
    const float delta = std::sqrt(ex * ex + ey * ey + ez * ez);
    if (dist >= size / theta + delta) {
      const float g = kG32 * c4[3] / (r2 * dist);
      ax += dx * g; ay += dy * g; az += dz * g;
      ++n_node;
      continue;
    }
    ++n_open;
    if (sp + 2 <= 64) {
      stack[sp++] = left[ni];
      stack[sp++] = right[ni];
    }
  }
  a3[0] = ax;
  a3[1] = ay;
  a3[2] = az;
  if (cnt) {
    cnt[0] = n_node;
    cnt[1] = n_pp;
    cnt[2] = n_open;
  }
}
}  // namespace

void or_bh_walk_f32(const float* sb, const uint32_t* order, int64_t n, int64_t k, const uint32_t* left,
                    const uint32_t* right, const float* com4, const float* bmin3, const float* bmax3, int64_t L,
                    float theta, float s2, float* acc) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t t = 0; t < n; ++t) bh_walk_one(sb, t, n, k, left, right, com4, bmin3, bmax3, L, theta, s2,
                                              acc + 3 * (int64_t)order[t], nullptr);
}

void or_bh_walk_count(const float* sb, int64_t n, int64_t k, const uint32_t* left, const uint32_t* right,
                      const float* com4, const float* bmin3, const float* bmax3, int64_t L, float theta, float s2,
                      const int64_t* targets, int64_t nt, int64_t* counts, float* acc3) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < nt; ++q) {
    float a3[3];
    bh_walk_one(sb, targets[q], n, k, left, right, com4, bmin3, bmax3, L, theta, s2, a3, counts + 3 * q);
    if (acc3)
      for (int c = 0; c < 3; ++c) acc3[3 * q + c] = a3[c];
  }
}

// ie/tools/gpu-bh-verify/src/main.rs:140-158 (FP64 direct softened sum), on selected targets.
void or_direct_f64(const double* xyz, const double* m, int64_t n, double soft, const int64_t* targets, int64_t nt,
                   double* acc, double* pot, int nthreads) {
  set_threads(nthreads);
  const double s2 = soft * soft;
  if (!targets) nt = n;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < nt; ++q) {
    const int64_t i = targets ? targets[q] : q;
    double a0 = 0, a1 = 0, a2 = 0, ph = 0;
    const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = xyz[3 * j] - xi, dy = xyz[3 * j + 1] - yi, dz = xyz[3 * j + 2] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + s2;
      const double rs = std::sqrt(r2);
      const double g = kG * m[j] / (r2 * rs);
      a0 += dx * g; a1 += dy * g; a2 += dz * g;
      ph -= kG * m[j] / rs;
    }
    acc[3 * q] = a0; acc[3 * q + 1] = a1; acc[3 * q + 2] = a2;
    if (pot) pot[q] = ph;
  }
}

void or_octree_f64(const double* xyz, const double* m, int64_t n, double theta, double soft, int leaf_cap,
                   double* acc) {
  Octree t(xyz, m, n, theta, soft, leaf_cap);
  t.accelerations(xyz, m, acc, 0);
}

int64_t or_grid_pairs(const double* xyz, int64_t n, double cell, double radius, int64_t* pairs, int64_t cap) {
  Grid g(xyz, n, cell, false);
  int64_t cnt = 0;
  const double r2max = radius * radius;
  g.for_each_pair(xyz, [&](int64_t i, int64_t j) {
    if (radius > 0) {
      const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                   dz = xyz[3 * i + 2] - xyz[3 * j + 2];
      if (dx * dx + dy * dy + dz * dz >= r2max) return;
    }
    if (cnt < cap) {
      pairs[2 * cnt] = i;
      pairs[2 * cnt + 1] = j;
    }
    ++cnt;
  });
  return cnt <= cap ? cnt : -cnt;
}

// ie/crates/engine/src/hydrostatic.rs:239-256 — rho_i = m_i W(0,h_i) + sum_{r<h_ij} m_j W(r,h_ij).
void or_density(const or_state* s, double* rho) {
  const int64_t n = s->n;
  double cell = 0.0;
  for (int64_t i = 0; i < n; ++i) cell = std::max(cell, s->h[i]);
  Grid g(s->xyz, n, cell, s->brute != 0);
  for (int64_t i = 0; i < n; ++i) rho[i] = s->m[i] * sph_w64(0.0, s->h[i]);
  g.for_each_pair(s->xyz, [&](int64_t i, int64_t j) {
    const double dx = s->xyz[3 * i] - s->xyz[3 * j], dy = s->xyz[3 * i + 1] - s->xyz[3 * j + 1],
                 dz = s->xyz[3 * i + 2] - s->xyz[3 * j + 2];
    const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double hij = 0.5 * (s->h[i] + s->h[j]);
    if (r < hij) {
      const double w = sph_w64(r, hij);
      rho[i] += s->m[j] * w;
      rho[j] += s->m[i] * w;
    }
  });
}

// ie/crates/engine/src/hydrostatic.rs:294-340 — gravity + symmetric SPH-EOS pressure + Monaghan AV
// + du/dt. Gravity per gravity_mode: octree (bhtree.rs, theta), FP64 direct, or none.
void or_forces(const or_state* s, const double* rho, double* acc, double* dudt) {
  const int64_t n = s->n;
  if (s->gravity_mode == 0) {
    Octree t(s->xyz, s->m, n, s->theta, s->softening, s->leaf_cap);
    t.accelerations(s->xyz, s->m, acc, s->nthreads);
  } else if (s->gravity_mode == 1) {
    or_direct_f64(s->xyz, s->m, n, s->softening, nullptr, n, acc, nullptr, s->nthreads);
  } else {
    std::fill(acc, acc + 3 * n, 0.0);
  }
  std::fill(dudt, dudt + n, 0.0);
  std::vector<double> p(n), c(n);
  for (int64_t i = 0; i < n; ++i) {
    const Till t(s->eos + 10 * s->mat[i]);
    p[i] = pressure64(t, rho[i], s->u[i]);
    c[i] = s->cs_fd ? sound_speed_wgsl64(t, rho[i], s->u[i]) : std::sqrt(sound_speed_sq64(t, rho[i], s->u[i]));
  }
  double cell = 0.0;
  for (int64_t i = 0; i < n; ++i) cell = std::max(cell, s->h[i]);
  Grid g(s->xyz, n, cell, s->brute != 0);
  g.for_each_pair(s->xyz, [&](int64_t i, int64_t j) {
    const double dp[3] = {s->xyz[3 * i] - s->xyz[3 * j], s->xyz[3 * i + 1] - s->xyz[3 * j + 1],
                          s->xyz[3 * i + 2] - s->xyz[3 * j + 2]};
    const double r = std::sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
    const double hij = 0.5 * (s->h[i] + s->h[j]);
    if (r >= hij || r < 1.0e-9) return;
    const double dv[3] = {s->vel[3 * i] - s->vel[3 * j], s->vel[3 * i + 1] - s->vel[3 * j + 1],
                          s->vel[3 * i + 2] - s->vel[3 * j + 2]};
    const double vr = dv[0] * dp[0] + dv[1] * dp[1] + dv[2] * dp[2];
    double pi_ij = 0.0;
    if (vr < 0.0) {
      const double mu = hij * vr / (r * r + 0.01 * hij * hij);
      const double c_bar = 0.5 * (c[i] + c[j]);
      const double rho_bar = 0.5 * (rho[i] + rho[j]);
      pi_ij = (-s->av_alpha * c_bar * mu + s->av_beta * mu * mu) / rho_bar;
    }
    const double coeff = p[i] / (rho[i] * rho[i]) + p[j] / (rho[j] * rho[j]) + pi_ij;
    const double dw = sph_dw64(r, hij);
    double grad[3];
    for (int q = 0; q < 3; ++q) grad[q] = dp[q] / r * dw;
    for (int q = 0; q < 3; ++q) {
      acc[3 * i + q] += grad[q] * (-coeff * s->m[j]);
      acc[3 * j + q] += grad[q] * (coeff * s->m[i]);
    }
    const double vdg = dv[0] * grad[0] + dv[1] * grad[1] + dv[2] * grad[2];
    dudt[i] += 0.5 * s->m[j] * coeff * vdg;
    dudt[j] += 0.5 * s->m[i] * coeff * vdg;
  });
}

// ie/crates/engine/src/hydrostatic.rs:366-374 (cs_fd=1: the WGSL cs_signal sound speed, sph_step.wgsl:263-269).
double or_courant_dt(const or_state* s, const double* rho, double cfl) {
  double dt_min = std::numeric_limits<double>::infinity();
  for (int64_t i = 0; i < s->n; ++i) {
    const Till t(s->eos + 10 * s->mat[i]);
    const double c = s->cs_fd ? sound_speed_wgsl64(t, rho[i], s->u[i]) : std::sqrt(sound_speed_sq64(t, rho[i], s->u[i]));
    const double v = std::sqrt(s->vel[3 * i] * s->vel[3 * i] + s->vel[3 * i + 1] * s->vel[3 * i + 1] +
                               s->vel[3 * i + 2] * s->vel[3 * i + 2]);
    dt_min = std::min(dt_min, s->h[i] / std::max(c + v, 1.0));
  }
  return cfl * dt_min;
}

// ie/crates/engine/src/hydrostatic.rs:345-359 (KDK with two force evaluations), with the
// impact-run adaptive dt (impact-run:403-410) when dt <= 0.
double or_kdk(or_state* s, double dt_fixed, double cfl, int32_t steps) {
  const int64_t n = s->n;
  std::vector<double> rho(n), acc(3 * n), dudt(n);
  double t = 0.0;
  // Optional Newton h-solve before each density pass; its candidate radius is 1.5x the largest current h.
  auto solve_h = [&]() {
    if (!s->h_solve) return;
    double hmax = 0.0;
    for (int64_t i = 0; i < n; ++i) hmax = std::max(hmax, s->h[i]);
    or_hsolve(s->xyz, s->m, n, s->h_eta, s->h_tol, 40, 1.5 * hmax, s->h, nullptr, nullptr, s->nthreads);
  };
  for (int32_t st = 0; st < steps; ++st) {
    solve_h();
    or_density(s, rho.data());
    const double dt = dt_fixed > 0 ? dt_fixed : or_courant_dt(s, rho.data(), cfl);
    or_forces(s, rho.data(), acc.data(), dudt.data());
    for (int64_t i = 0; i < n; ++i) {
      for (int q = 0; q < 3; ++q) s->vel[3 * i + q] += acc[3 * i + q] * (0.5 * dt);
      s->u[i] = std::max(s->u[i] + dudt[i] * (0.5 * dt), 0.0);
      for (int q = 0; q < 3; ++q) s->xyz[3 * i + q] += s->vel[3 * i + q] * dt;
    }
    solve_h();
    or_density(s, rho.data());
    or_forces(s, rho.data(), acc.data(), dudt.data());
    for (int64_t i = 0; i < n; ++i) {
      for (int q = 0; q < 3; ++q) s->vel[3 * i + q] += acc[3 * i + q] * (0.5 * dt);
      s->u[i] = std::max(s->u[i] + dudt[i] * (0.5 * dt), 0.0);
    }
    t += dt;
  }
  return t;
}

// ie/crates/engine/src/hydrostatic.rs:283-291 (relax_step) over `steps` steps: density, then the pressure-only
// accelerations of hydrostatic.rs:258-281 (gravity + the symmetric SPH-EOS pressure term, no artificial
// viscosity), then v = (v + a dt) damp, x += v dt. u is left unchanged (isothermal relaxation).
void or_relax(or_state* s, double dt, double damp, int32_t steps) {
  const int64_t n = s->n;
  std::vector<double> rho(n), acc(3 * n), dudt(n);
  const double alpha = s->av_alpha, beta = s->av_beta;
  s->av_alpha = 0.0;
  s->av_beta = 0.0;
  for (int32_t st = 0; st < steps; ++st) {
    or_density(s, rho.data());
    or_forces(s, rho.data(), acc.data(), dudt.data());
    for (int64_t i = 0; i < n; ++i)
      for (int q = 0; q < 3; ++q) {
        s->vel[3 * i + q] = (s->vel[3 * i + q] + acc[3 * i + q] * dt) * damp;
        s->xyz[3 * i + q] += s->vel[3 * i + q] * dt;
      }
  }
  s->av_alpha = alpha;
  s->av_beta = beta;
}

// ie/crates/engine/src/gpu_sph.rs:628-649 (KE, IE, direct softened PE) plus momentum and angular
// momentum about the origin (impact-run:528-536 takes L_z about the COM; callers that want that
// pass COM-frame states).
void or_energy(const or_state* s, double* out) {
  const int64_t n = s->n;
  double ke = 0, ie = 0, p[3] = {0, 0, 0}, L[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    const double* v = s->vel + 3 * i;
    const double* x = s->xyz + 3 * i;
    ke += 0.5 * s->m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    ie += s->m[i] * s->u[i];
    for (int q = 0; q < 3; ++q) p[q] += s->m[i] * v[q];
    L[0] += s->m[i] * (x[1] * v[2] - x[2] * v[1]);
    L[1] += s->m[i] * (x[2] * v[0] - x[0] * v[2]);
    L[2] += s->m[i] * (x[0] * v[1] - x[1] * v[0]);
  }
  std::vector<double> pot(n);
  std::vector<double> acc(3 * n);
  or_direct_f64(s->xyz, s->m, n, s->softening, nullptr, n, acc.data(), pot.data(), s->nthreads);
  double pe = 0;
  for (int64_t i = 0; i < n; ++i) pe += 0.5 * s->m[i] * pot[i];
  out[0] = ke; out[1] = ie; out[2] = pe;
  out[3] = p[0]; out[4] = p[1]; out[5] = p[2];
  out[6] = L[0]; out[7] = L[1]; out[8] = L[2];
}

// NEW — no reference counterpart (the reference keeps h fixed, hydrostatic.rs:58-61). Safeguarded
// Newton on f(h) = rho~(h) - m_i (eta/h)^3 with a bracket [lo, hi]: f < 0 below the root, > 0 above.
int64_t or_hsolve(const double* xyz, const double* m, int64_t n, double eta, double tol, int32_t max_iter,
                  double h_max, double* h, int32_t* iters, double* resid, int nthreads) {
  set_threads(nthreads);
  std::vector<std::vector<std::pair<double, int64_t>>> cand;
  candidates(xyz, n, h_max, cand);
  const double eta3 = eta * eta * eta;
  int64_t bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
  for (int64_t i = 0; i < n; ++i) {
    double hh = std::min(std::max(h[i], 1e-30), h_max);
    double lo = 0.0, hi = std::numeric_limits<double>::infinity();
    int32_t it = 0;
    double res = 0;
    bool ok = false;
    for (;;) {
      double rho, drho;
      own_density(cand[i], m, i, hh, rho, drho);
      const double target = m[i] * eta3 / (hh * hh * hh);
      const double f = rho - target;
      res = std::fabs(f) / rho;
      if (res <= tol) { ok = true; break; }
      if (it >= max_iter) break;
      if (f < 0) lo = hh; else hi = hh;
      if (f < 0 && hh >= h_max) break;  // isolated: no root below h_max
      const double df = drho + 3.0 * target / hh;
      double hn = (df > 0) ? hh - f / df : -1.0;
      if (!(hn > lo && hn < hi)) {
        if (std::isinf(hi)) hn = 2.0 * hh;
        else hn = 0.5 * (lo + hi);
      }
      hh = std::min(hn, h_max);
      ++it;
    }
    h[i] = hh;
    if (iters) iters[i] = it;
    if (resid) resid[i] = res;
    if (!ok) ++bad;
  }
  return bad;
}

void or_hresidual(const double* xyz, const double* m, int64_t n, double eta, const double* h, double* resid,
                  int32_t* nngb, int nthreads) {
  set_threads(nthreads);
  double hmax = 0;
  for (int64_t i = 0; i < n; ++i) hmax = std::max(hmax, h[i]);
  std::vector<std::vector<std::pair<double, int64_t>>> cand;
  candidates(xyz, n, hmax * 1.000001, cand);
  const double eta3 = eta * eta * eta;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    double rho, drho;
    own_density(cand[i], m, i, h[i], rho, drho);
    const double f = rho - m[i] * eta3 / (h[i] * h[i] * h[i]);
    resid[i] = std::fabs(f) / rho;
    if (nngb) {
      int32_t c = 0;
      for (const auto& x : cand[i])
        if (x.first < h[i]) ++c;
      nngb[i] = c;
    }
  }
}

}  // extern "C"
