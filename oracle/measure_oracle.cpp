// CPU oracle for the disk / remnant measurement and accretion clumps — TEST INFRASTRUCTURE ONLY.
//
// A plain restatement of the reference's Rust f64 code, in its own operation order (compiled with
// -ffp-contract=off, no FMA), used as the checker for the HIP measurement. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it; the product never does.
//
//   or_sum_oi          ie/tools/impact-run/src/main.rs:195-205 (magnitude-sorted Kahan sum)
//   or_measure_ir      ie/tools/impact-run/src/main.rs:207-215, 232-247, 732-778 (com, perigee, measure)
//                      + main.rs:802-878 (moon_candidate: O(k^2) union-find over the disk)
//   or_engine_frame    ie/crates/engine/src/gpu_sph.rs:399-453 / 478-500 (the preamble disk_stats_json,
//                      moonlet_bodies, largest_moonlet_orbit and disk_moonlets share) with orbit.rs:91-116
//   or_find_clumps     ie/crates/engine/src/accretion.rs:73-163 over neighbors.rs:41-85 (NeighborGrid pairs,
//                      brute force below 512 bodies)
//   or_body_bulk       ie/crates/engine/src/gpu_sph.rs:279-316 (clipped centre, 98%-mass radius)
//   or_assemble_n      ie/crates/engine/src/gpu_sph.rs:318-339 (placement of relaxed bodies)
//
// Two places where the reference is not deterministic are fixed here, the same way the HIP library fixes them:
// clumps are visited in order of their lowest member (the reference iterates a HashMap), so the Moon candidate
// among clumps of equal mass is the one with the lowest member index.
#include "measure_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace {

constexpr double kG = 6.674e-11;  // main.rs:26, orbit.rs:18
constexpr double kPI = 3.14159265358979323846;
constexpr double kFourThirdsPi = 4.0 / 3.0 * kPI;  // accretion.rs:26

// main.rs:195-205.
double sum_oi(std::vector<double>& t) {
  std::stable_sort(t.begin(), t.end(), [](double a, double b) { return std::fabs(a) < std::fabs(b); });
  double s = 0.0, c = 0.0;
  for (double x : t) {
    const double y = x - c;
    const double z = s + y;
    c = (z - s) - y;
    s = z;
  }
  return s;
}

// main.rs:232-247.
bool perigee_ir(const double p[3], const double v[3], double mu, double* pg) {
  const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  if (r == 0.0) {
    *pg = 0.0;
    return true;
  }
  const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double energy = 0.5 * v2 - mu / r;
  if (energy >= 0.0) return false;
  const double a = -mu / (2.0 * energy);
  const double h[3] = {p[1] * v[2] - p[2] * v[1], p[2] * v[0] - p[0] * v[2], p[0] * v[1] - p[1] * v[0]};
  const double h2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
  const double e = std::sqrt(std::fmax(1.0 + 2.0 * energy * h2 / (mu * mu), 0.0));
  *pg = a * (1.0 - e);
  return true;
}

// orbit.rs:91-94 (glam: length = sqrt(x x + y y + z z)).
bool is_bound(const double p[3], const double v[3], double mu) {
  const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  return r > 0.0 && 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - mu / r < 0.0;
}

// orbit.rs:96-116 (glam cross: (y rz - ry z, z rx - rz x, x ry - rx y)).
bool perigee_orbit(const double p[3], const double v[3], double mu, double* pg) {
  const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  if (r == 0.0) {
    *pg = 0.0;
    return true;
  }
  const double energy = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - mu / r;
  const double c[3] = {p[1] * v[2] - v[1] * p[2], p[2] * v[0] - v[2] * p[0], p[0] * v[1] - v[0] * p[1]};
  const double h = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const double e = std::sqrt(std::fmax(1.0 + 2.0 * energy * h * h / (mu * mu), 0.0));
  if (energy >= 0.0) {
    if (e <= 1.0 || h == 0.0) return false;
    *pg = h * h / mu / (1.0 + e);
    return true;
  }
  const double a = -mu / (2.0 * energy);
  *pg = a * (1.0 - e);
  return true;
}

// Union-find with path compression (main.rs:817-824, accretion.rs:90-103).
size_t uf_find(std::vector<size_t>& p, size_t i) {
  size_t r = i;
  while (p[r] != r) r = p[r];
  size_t c = i;
  while (p[c] != r) {
    const size_t nx = p[c];
    p[c] = r;
    c = nx;
  }
  return r;
}

// Members per component, components in order of their lowest member, members ascending.
std::vector<std::vector<size_t>> components(std::vector<size_t>& parent) {
  const size_t k = parent.size();
  std::vector<std::vector<size_t>> out;
  std::unordered_map<size_t, size_t> slot;
  for (size_t a = 0; a < k; ++a) {
    const size_t r = uf_find(parent, a);
    auto it = slot.find(r);
    if (it == slot.end()) {
      slot.emplace(r, out.size());
      out.push_back({a});
    } else {
      out[it->second].push_back(a);
    }
  }
  return out;
}

// The shared remnant preamble: radius about com enclosing 85% of m_total, sequential running sum in radius
// order (stable), defaults to the largest radius and m_total (main.rs:738-747, gpu_sph.rs:411-420).
void remnant(const or_particle* P, int64_t n, const double com[3], double m_total, double* r_rem, double* m_rem) {
  std::vector<std::pair<double, double>> radii(n);
  for (int64_t i = 0; i < n; ++i) {
    const double d[3] = {(double)P[i].pos[0] - com[0], (double)P[i].pos[1] - com[1], (double)P[i].pos[2] - com[2]};
    radii[i] = {std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), (double)P[i].mass};
  }
  std::stable_sort(radii.begin(), radii.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  double cum = 0.0;
  *r_rem = n > 0 ? radii.back().first : 0.0;
  *m_rem = m_total;
  for (const auto& rm : radii) {
    cum += rm.second;
    if (cum >= 0.85 * m_total) {
      *r_rem = rm.first;
      *m_rem = cum;
      break;
    }
  }
}

}  // namespace

extern "C" {

double or_sum_oi(const double* terms, int64_t n) {
  std::vector<double> t(terms, terms + n);
  return sum_oi(t);
}

int or_perigee(int32_t which, const double* rel_p, const double* rel_v, double mu, double* pg) {
  return which == 0 ? (perigee_ir(rel_p, rel_v, mu, pg) ? 1 : 0) : (perigee_orbit(rel_p, rel_v, mu, pg) ? 1 : 0);
}

void or_measure_ir(const or_particle* P, int64_t n, int64_t n_earth, or_measure* out, int32_t* cls) {
  std::memset(out, 0, sizeof *out);
  std::vector<double> masses(n);
  for (int64_t i = 0; i < n; ++i) masses[i] = (double)P[i].mass;
  const double m_total = sum_oi(masses);
  // com (main.rs:207-215): sequential, divided by its own sequential mass sum.
  double c[3] = {0.0, 0.0, 0.0}, mc = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    for (int q = 0; q < 3; ++q) c[q] += (double)P[i].pos[q] * (double)P[i].mass;
    mc += (double)P[i].mass;
  }
  const double com[3] = {c[0] / mc, c[1] / mc, c[2] / mc};
  std::vector<double> vx(n), vy(n), vz(n);
  for (int64_t i = 0; i < n; ++i) {
    vx[i] = (double)P[i].vel[0] * (double)P[i].mass;
    vy[i] = (double)P[i].vel[1] * (double)P[i].mass;
    vz[i] = (double)P[i].vel[2] * (double)P[i].mass;
  }
  const double v_com[3] = {sum_oi(vx) / m_total, sum_oi(vy) / m_total, sum_oi(vz) / m_total};
  double r_remnant, m_remnant;
  remnant(P, n, com, m_total, &r_remnant, &m_remnant);
  const double mu = kG * m_remnant;
  std::vector<double> e_disk, t_disk, e_esc, t_esc, e_rem, t_rem;
  std::vector<size_t> disk_idx;
  for (int64_t i = 0; i < n; ++i) {
    const double rp[3] = {(double)P[i].pos[0] - com[0], (double)P[i].pos[1] - com[1], (double)P[i].pos[2] - com[2]};
    const double rv[3] = {(double)P[i].vel[0] - v_com[0], (double)P[i].vel[1] - v_com[1],
                          (double)P[i].vel[2] - v_com[2]};
    const bool earth = i < n_earth;
    const double m = (double)P[i].mass;
    double pg;
    int k;
    if (!perigee_ir(rp, rv, mu, &pg)) {
      (earth ? e_esc : t_esc).push_back(m);
      k = 2;
    } else if (pg > r_remnant) {
      (earth ? e_disk : t_disk).push_back(m);
      disk_idx.push_back((size_t)i);
      k = 1;
    } else {
      (earth ? e_rem : t_rem).push_back(m);
      k = 0;
    }
    if (cls) cls[i] = k;
  }
  const double ed = sum_oi(e_disk), td = sum_oi(t_disk);
  const double disk = ed + td;
  out->earth_frac = disk > 0.0 ? 100.0 * ed / disk : 0.0;
  out->disk_kg = disk;
  out->e_disk = ed;
  out->t_disk = td;
  out->e_rem = sum_oi(e_rem);
  out->t_rem = sum_oi(t_rem);
  out->e_esc = sum_oi(e_esc);
  out->t_esc = sum_oi(t_esc);
  out->r_remnant = r_remnant;
  out->m_total = m_total;
  out->m_remnant = m_remnant;
  for (int q = 0; q < 3; ++q) {
    out->com[q] = com[q];
    out->v_com[q] = v_com[q];
  }
  out->n_disk = (int64_t)disk_idx.size();
  // moon_candidate (main.rs:802-878).
  const size_t k = disk_idx.size();
  if (k < 2) return;
  double sh = 0.0;
  for (size_t i : disk_idx) sh += (double)P[i].h;
  const double link = 2.0 * (sh / (double)k);
  const double link2 = link * link;
  out->link = link;
  std::vector<size_t> parent(k);
  std::iota(parent.begin(), parent.end(), 0);
  for (size_t a = 0; a < k; ++a)
    for (size_t b = a + 1; b < k; ++b) {
      const float* pa = P[disk_idx[a]].pos;
      const float* pb = P[disk_idx[b]].pos;
      const double dx = (double)(pa[0] - pb[0]), dy = (double)(pa[1] - pb[1]), dz = (double)(pa[2] - pb[2]);
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 <= link2) {
        const size_t ra = uf_find(parent, a), rb = uf_find(parent, b);
        if (ra != rb) parent[ra] = rb;
      }
    }
  const auto groups = components(parent);
  const double rho_rem = m_remnant / (4.0 / 3.0 * kPI * (r_remnant * r_remnant * r_remnant));
  double best_m = 0.0, best_e = 0.0;
  size_t best_n = 0, n_bound = 0, n_clumps = 0;
  bool found = false;
  for (const auto& g : groups) {
    if (g.size() < 2) continue;
    ++n_clumps;
    double m = 0.0;
    for (size_t a : g) m += (double)P[disk_idx[a]].mass;
    double cv[3] = {0, 0, 0}, cp[3] = {0, 0, 0}, vol = 0.0;
    for (size_t a : g) {
      const or_particle& p = P[disk_idx[a]];
      for (int x = 0; x < 3; ++x) {
        cv[x] += (double)p.vel[x] * (double)p.mass;
        cp[x] += (double)p.pos[x] * (double)p.mass;
      }
      vol += (double)p.mass / (double)p.rho;
    }
    for (int x = 0; x < 3; ++x) {
      cv[x] /= m;
      cp[x] /= m;
    }
    const double clump_rho = vol > 0.0 ? m / vol : 2700.0;
    double ke = 0.0;
    for (size_t a : g) {
      const or_particle& p = P[disk_idx[a]];
      const double d0 = (double)p.vel[0] - cv[0], d1 = (double)p.vel[1] - cv[1], d2 = (double)p.vel[2] - cv[2];
      ke += 0.5 * (double)p.mass * (d0 * d0 + d1 * d1 + d2 * d2);
    }
    double pe = 0.0;
    for (size_t a = 0; a < g.size(); ++a)
      for (size_t b = a + 1; b < g.size(); ++b) {
        const or_particle& A = P[disk_idx[g[a]]];
        const or_particle& Bp = P[disk_idx[g[b]]];
        const double d[3] = {(double)(A.pos[0] - Bp.pos[0]), (double)(A.pos[1] - Bp.pos[1]),
                             (double)(A.pos[2] - Bp.pos[2])};
        const double r = std::fmax(std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), 1.0);
        pe -= kG * (double)A.mass * (double)Bp.mass / r;
      }
    const bool bound = ke + pe < 0.0;
    const double d_roche = 2.44 * r_remnant * std::cbrt(rho_rem / clump_rho);
    const double dist = std::sqrt((cp[0] - com[0]) * (cp[0] - com[0]) + (cp[1] - com[1]) * (cp[1] - com[1]) +
                                  (cp[2] - com[2]) * (cp[2] - com[2]));
    if (bound) ++n_bound;
    if (bound && dist > d_roche && m > best_m) {
      best_m = m;
      best_e = 0.0;
      for (size_t a : g)
        if ((int64_t)disk_idx[a] < n_earth) best_e += (double)P[disk_idx[a]].mass;
      best_n = g.size();
      found = true;
    }
  }
  out->n_clumps = (int64_t)n_clumps;
  out->n_bound = (int64_t)n_bound;
  if (found) {
    out->clump_kg = best_m;
    out->clump_earth_frac = 100.0 * best_e / best_m;
    out->clump_n = (int64_t)best_n;
  }
}

void or_engine_frame(const or_particle* P, int64_t n, int32_t require_bound, double* frame, double* esc_e_t,
                     int32_t* cls) {
  // gpu_sph.rs:406-410: sequential sums.
  double m_total = 0.0;
  for (int64_t i = 0; i < n; ++i) m_total += (double)P[i].mass;
  double sp[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i)
    for (int q = 0; q < 3; ++q) {
      sp[q] += (double)P[i].pos[q] * (double)P[i].mass;
      sv[q] += (double)P[i].vel[q] * (double)P[i].mass;
    }
  const double com[3] = {sp[0] / m_total, sp[1] / m_total, sp[2] / m_total};
  const double v_com[3] = {sv[0] / m_total, sv[1] / m_total, sv[2] / m_total};
  double r_remnant, m_remnant;
  remnant(P, n, com, m_total, &r_remnant, &m_remnant);
  const double mu = kG * m_remnant;
  double e_disk = 0.0, t_disk = 0.0, esc = 0.0, sum_h = 0.0;
  for (int64_t i = 0; i < n; ++i) sum_h += (double)P[i].h;
  for (int64_t i = 0; i < n; ++i) {
    const double rp[3] = {(double)P[i].pos[0] - com[0], (double)P[i].pos[1] - com[1], (double)P[i].pos[2] - com[2]};
    const double rv[3] = {(double)P[i].vel[0] - v_com[0], (double)P[i].vel[1] - v_com[1],
                          (double)P[i].vel[2] - v_com[2]};
    const double m = (double)P[i].mass;
    const bool bound = is_bound(rp, rv, mu);
    double pg;
    const bool disk = (!require_bound || bound) && perigee_orbit(rp, rv, mu, &pg) && pg > r_remnant;
    if (require_bound && !bound) esc += m;  // gpu_sph.rs:429-432
    if (disk) {
      if (P[i].prov == 0u) e_disk += m;
      else t_disk += m;
    }
    if (cls) cls[i] = disk ? 1 : (bound ? 0 : 2);
  }
  frame[0] = m_total;
  for (int q = 0; q < 3; ++q) {
    frame[1 + q] = com[q];
    frame[4 + q] = v_com[q];
  }
  frame[7] = r_remnant;
  frame[8] = m_remnant;
  frame[9] = mu;
  frame[10] = e_disk + t_disk;
  frame[11] = 2.0 * (sum_h / (double)n);  // gpu_sph.rs:457-458
  if (esc_e_t) {
    esc_e_t[0] = e_disk;
    esc_e_t[1] = t_disk;
    esc_e_t[2] = esc;
  }
}

int64_t or_find_clumps(int64_t n, const double* pos3, const double* vel3, const double* mass, const double* rho,
                       double linking_length, double g, double softening, const double* central3,
                       double central_mass, double central_radius, or_clump* out, int64_t cap, int64_t* label) {
  if (n <= 0) return 0;
  std::vector<size_t> parent((size_t)n);
  std::iota(parent.begin(), parent.end(), 0);
  const double ll2 = linking_length * linking_length;
  auto link = [&](int64_t i, int64_t j) {
    const double d[3] = {pos3[3 * i] - pos3[3 * j], pos3[3 * i + 1] - pos3[3 * j + 1],
                         pos3[3 * i + 2] - pos3[3 * j + 2]};
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= ll2) {
      const size_t a = uf_find(parent, (size_t)i), b = uf_find(parent, (size_t)j);
      if (a != b) parent[a] = b;
    }
  };
  // neighbors.rs:41-85: brute force below 512 bodies, else a hash of cells of side max(link, 1e-9).
  if (n < 512) {
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = i + 1; j < n; ++j) link(i, j);
  } else {
    const double cell = std::max(linking_length, 1.0e-9);
    // `as i32` saturates (NaN -> 0); neighbour offsets are taken in 64 bits.
    auto sat = [](double v) -> int64_t {
      if (std::isnan(v)) return 0;
      return (int64_t)std::min(2147483647.0, std::max(-2147483648.0, v));
    };
    auto key = [&](int64_t i) {
      return std::make_tuple(sat(std::floor(pos3[3 * i] / cell)), sat(std::floor(pos3[3 * i + 1] / cell)),
                             sat(std::floor(pos3[3 * i + 2] / cell)));
    };
    std::map<std::tuple<int64_t, int64_t, int64_t>, std::vector<int64_t>> cells;
    for (int64_t i = 0; i < n; ++i) cells[key(i)].push_back(i);
    for (int64_t i = 0; i < n; ++i) {
      const auto [cx, cy, cz] = key(i);
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            auto it = cells.find(std::make_tuple(cx + dx, cy + dy, cz + (int64_t)dz));
            if (it == cells.end()) continue;
            for (int64_t j : it->second)
              if (j > i) link(i, j);
          }
    }
  }
  const auto groups = components(parent);
  const double s2 = softening * softening;
  const double rho_c = central_mass / (kFourThirdsPi * (central_radius * central_radius * central_radius));
  int64_t kept = 0;
  for (const auto& mem : groups) {
    double m = 0.0;
    for (size_t i : mem) m += mass[i];
    if (!(m > 0.0)) {
      if (label)
        for (size_t i : mem) label[i] = -1;
      continue;
    }
    double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    for (size_t i : mem)
      for (int q = 0; q < 3; ++q) cp[q] += pos3[3 * i + q] * mass[i];
    for (size_t i : mem)
      for (int q = 0; q < 3; ++q) cv[q] += vel3[3 * i + q] * mass[i];
    for (int q = 0; q < 3; ++q) {
      cp[q] /= m;
      cv[q] /= m;
    }
    double vol = 0.0;
    for (size_t i : mem) vol += mass[i] / rho[i];
    const double clump_rho = vol > 0.0 ? m / vol : rho[mem[0]];
    const double radius = std::cbrt(m / (kFourThirdsPi * clump_rho));
    double ke = 0.0;
    for (size_t i : mem) {
      const double d[3] = {vel3[3 * i] - cv[0], vel3[3 * i + 1] - cv[1], vel3[3 * i + 2] - cv[2]};
      ke += 0.5 * mass[i] * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    double pe = 0.0;
    for (size_t a = 0; a < mem.size(); ++a)
      for (size_t b = a + 1; b < mem.size(); ++b) {
        const size_t ia = mem[a], ib = mem[b];
        const double d[3] = {pos3[3 * ia] - pos3[3 * ib], pos3[3 * ia + 1] - pos3[3 * ib + 1],
                             pos3[3 * ia + 2] - pos3[3 * ib + 2]};
        const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + s2);
        pe -= g * mass[ia] * mass[ib] / r;
      }
    const double d_roche = 2.44 * central_radius * std::cbrt(rho_c / clump_rho);
    const double e[3] = {cp[0] - central3[0], cp[1] - central3[1], cp[2] - central3[2]};
    if (kept < cap) {
      or_clump& c = out[kept];
      c.first = (int64_t)mem[0];
      c.n_members = (int64_t)mem.size();
      c.mass = m;
      for (int q = 0; q < 3; ++q) {
        c.com_pos[q] = cp[q];
        c.com_vel[q] = cv[q];
      }
      c.rho = clump_rho;
      c.radius = radius;
      c.internal_ke = ke;
      c.self_pe = pe;
      c.earth_mass = 0.0;
      c.bound = ke + pe < 0.0 ? 1 : 0;
      c.outside_roche = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) > d_roche ? 1 : 0;
    }
    if (label)
      for (size_t i : mem) label[i] = kept;
    ++kept;
  }
  return kept;
}

}  // extern "C"

namespace {
// gpu_sph.rs:283-304: com and bulk_radius over a selection.
void bulk_com(const or_particle* P, const std::vector<int64_t>& sel, double c[3]) {
  double mm = 0.0;
  for (int64_t i : sel) mm += (double)P[i].mass;
  c[0] = c[1] = c[2] = 0.0;
  if (mm <= 0.0) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int64_t i : sel)
    for (int q = 0; q < 3; ++q) s[q] += (double)P[i].pos[q] * (double)P[i].mass;
  for (int q = 0; q < 3; ++q) c[q] = s[q] / mm;
}
double radius_about(const or_particle& p, const double c[3]) {
  const double d[3] = {(double)p.pos[0] - c[0], (double)p.pos[1] - c[1], (double)p.pos[2] - c[2]};
  return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}
double bulk_radius(const or_particle* P, const std::vector<int64_t>& sel, const double c[3]) {
  double mm = 0.0;
  for (int64_t i : sel) mm += (double)P[i].mass;
  std::vector<std::pair<double, double>> rr;
  rr.reserve(sel.size());
  for (int64_t i : sel) rr.push_back({radius_about(P[i], c), (double)P[i].mass});
  std::stable_sort(rr.begin(), rr.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  double cum = 0.0, r = 0.0;
  for (const auto& x : rr) {
    cum += x.second;
    r = x.first;
    if (cum >= 0.98 * mm) break;
  }
  return r;
}
}  // namespace

extern "C" {

void or_body_bulk(const or_particle* P, int64_t n, uint32_t prov, double* out5) {
  std::vector<int64_t> ps;
  for (int64_t i = 0; i < n; ++i)
    if (P[i].prov == prov) ps.push_back(i);
  double m = 0.0;
  for (int64_t i : ps) m += (double)P[i].mass;
  double c[3];
  bulk_com(P, ps, c);
  double r = bulk_radius(P, ps, c);
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int64_t> keep;
    for (int64_t i : ps)
      if (radius_about(P[i], c) <= 1.5 * r) keep.push_back(i);
    if (keep.empty()) break;
    bulk_com(P, keep, c);
    r = bulk_radius(P, keep, c);
  }
  out5[0] = c[0];
  out5[1] = c[1];
  out5[2] = c[2];
  out5[3] = m;
  out5[4] = r;
}

int64_t or_assemble_n(const or_particle* P, int64_t n, const double* placements, int32_t nbody, or_particle* out,
                      float* softening, float* dt) {
  std::vector<double> coms(3 * (size_t)nbody);
  for (int k = 0; k < nbody; ++k) {
    double b[5];
    or_body_bulk(P, n, (uint32_t)k, b);
    for (int q = 0; q < 3; ++q) coms[3 * k + q] = b[q];
  }
  int64_t kept = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (P[i].prov >= (uint32_t)nbody) continue;
    const double* pl = placements + 9 * P[i].prov;
    const double* c = &coms[3 * P[i].prov];
    const double l[3] = {(double)P[i].pos[0] - c[0], (double)P[i].pos[1] - c[1], (double)P[i].pos[2] - c[2]};
    const double sp[3] = {pl[6], pl[7], pl[8]};
    const double cr[3] = {sp[1] * l[2] - l[1] * sp[2], sp[2] * l[0] - l[2] * sp[0], sp[0] * l[1] - l[0] * sp[1]};
    or_particle q = P[i];
    for (int a = 0; a < 3; ++a) {
      q.pos[a] = (float)(l[a] + pl[a]);
      q.vel[a] = (float)(pl[3 + a] + cr[a]);
    }
    out[kept++] = q;
  }
  float min_h = INFINITY;
  for (int64_t i = 0; i < kept; ++i) min_h = std::fmin(min_h, out[i].h);
  *softening = 0.25f * min_h;
  double max_speed = 0.0;
  for (int k = 0; k < nbody; ++k) {
    const double* v = placements + 9 * k + 3;
    max_speed = std::fmax(max_speed, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
  }
  *dt = (float)(0.01 * (double)min_h / (20000.0 + max_speed));
  return kept;
}

}  // extern "C"
