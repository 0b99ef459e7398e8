// Seeded synthetic AMOEBA-style systems: rigid water boxes and water with a
// compact solute blob.  Host-only code; no device use.
//
// Water recipe (AMOEBA-03 water geometry and polarizabilities,
// GPU/params/water03.prm:135-156): O-H 0.9572 A, H-O-H 104.52 deg,
// q_O = -0.51966 e, q_H = +0.25983 e, local-frame dipoles 0.14 e.A on O
// along the bisector (towards the H atoms) and 0.02 e.A on each H along
// O->H, rotated into the global frame here; traceless quadrupoles with five
// independent N(0, 0.1) components stored directly in the internal rpole
// convention; alpha_O = 0.837, alpha_H = 0.496 A^3, Thole a = 0.39.  Each
// molecule is one polarization group, so its three intramolecular pairs get
// d = p = 0 and u = 1 (v1.2/source/initprm.f:417-466).
// Solute atoms: q ~ N(0, 0.3) neutralised, d ~ N(0, 0.1) per component,
// Q ~ N(0, 0.1) traceless, alpha ~ U(1, 1.75), a = 0.39, no exclusions.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/thp.h"

namespace {

// xoshiro256** seeded through splitmix64: identical streams on every host.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t t = z;
      t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
      t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
      s[i] = t ^ (t >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  double uniform() { return double(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double normal(double sd) {
    // Box-Muller, cosine branch only
    double u1 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = uniform();
    return sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  uint64_t below(uint64_t n) { return next() % n; }
};

constexpr double kBond = 0.9572;
constexpr double kAngleDeg = 104.52;
constexpr double kQO = -0.51966, kQH = 0.25983;
constexpr double kDipO = 0.14, kDipH = 0.02;
constexpr double kPolO = 0.837, kPolH = 0.496;
constexpr double kThole = 0.39;
constexpr double kQuadSd = 0.1;

// uniform random rotation (Shoemake 1992, "Uniform random rotations")
void random_rotation(Rng& g, double R[3][3]) {
  const double u1 = g.uniform(), u2 = g.uniform(), u3 = g.uniform();
  const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
  const double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
  const double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
  R[0][0] = 1 - 2 * (qy * qy + qz * qz);
  R[0][1] = 2 * (qx * qy - qz * qw);
  R[0][2] = 2 * (qx * qz + qy * qw);
  R[1][0] = 2 * (qx * qy + qz * qw);
  R[1][1] = 1 - 2 * (qx * qx + qz * qz);
  R[1][2] = 2 * (qy * qz - qx * qw);
  R[2][0] = 2 * (qx * qz - qy * qw);
  R[2][1] = 2 * (qy * qz + qx * qw);
  R[2][2] = 1 - 2 * (qx * qx + qy * qy);
}

void random_quadrupole(Rng& g, double* rp) {
  const double qxx = g.normal(kQuadSd), qyy = g.normal(kQuadSd);
  const double qxy = g.normal(kQuadSd), qxz = g.normal(kQuadSd), qyz = g.normal(kQuadSd);
  const double qzz = -qxx - qyy;
  rp[4] = qxx;
  rp[5] = qxy;
  rp[6] = qxz;
  rp[7] = qxy;
  rp[8] = qyy;
  rp[9] = qyz;
  rp[10] = qxz;
  rp[11] = qyz;
  rp[12] = qzz;
}

// lattice sites of an n_side^3 grid over [0,L)^3, optionally excluding a
// sphere, in a seeded random order
std::vector<int> shuffled_sites(Rng& g, int n_side, double L, double excl_r, double cx,
                                double cy, double cz) {
  std::vector<int> sites;
  const double a = L / n_side;
  for (int k = 0; k < n_side; ++k)
    for (int j = 0; j < n_side; ++j)
      for (int i = 0; i < n_side; ++i) {
        if (excl_r > 0) {
          const double dx = (i + 0.5) * a - cx, dy = (j + 0.5) * a - cy, dz = (k + 0.5) * a - cz;
          if (dx * dx + dy * dy + dz * dz < excl_r * excl_r) continue;
        }
        sites.push_back((k * n_side + j) * n_side + i);
      }
  for (size_t m = sites.size(); m > 1; --m) std::swap(sites[m - 1], sites[g.below(m)]);
  return sites;
}

}  // namespace

extern "C" int thp_pick_grid(double minsize) {
  // v1.2/source/kewald.f:67-73 (2,3,5-smooth even sizes)
  static const int multi[] = {2,   4,   6,   8,   10,  12,  16,  18,  20,  24,  30,  32,  36,
                              40,  48,  50,  54,  60,  64,  72,  80,  90,  96,  100, 108, 120,
                              128, 144, 150, 160, 162, 180, 192, 200, 216, 240, 250, 256, 270,
                              288, 300, 320, 324, 360, 384, 400, 432, 450, 480, 486, 500, 512,
                              540, 576, 600, 640, 648, 720, 750, 768, 800, 810, 864};
  for (int m : multi)
    if (m >= minsize - 1e-9) return m;
  return multi[sizeof(multi) / sizeof(multi[0]) - 1];
}

extern "C" thp_status thp_make_system(int n, double L, int n_solute_req, double blob_radius,
                                      uint64_t seed, double* x, double* y, double* z, double* rpole,
                                      double* polarity, double* thole, int* nscaled, int* scaled_ik,
                                      double* scaled_dpu) {
  if (n <= 0 || L <= 0 || n_solute_req < 0 || n_solute_req > n) return THP_ERR_ARG;
  Rng g(seed);
  // water molecules take multiples of 3; the solute absorbs the remainder
  const int nmol = (n - n_solute_req) / 3;
  const int nsol = n - 3 * nmol;
  if (nsol > 0 && blob_radius <= 0) return THP_ERR_ARG;
  const double cx = 0.5 * L, cy = 0.5 * L, cz = 0.5 * L;
  const double half = 0.5 * kAngleDeg * M_PI / 180.0;
  // molecular frame: bisector along +z, molecule in the xz plane
  const double hloc[2][3] = {{kBond * std::sin(half), 0.0, kBond * std::cos(half)},
                             {-kBond * std::sin(half), 0.0, kBond * std::cos(half)}};
  int nsc = 0;
  // ---- water on a jittered lattice (outside the blob when there is one)
  if (nmol > 0) {
    const double excl = nsol > 0 ? blob_radius + 1.5 : 0.0;
    const double vol_frac = nsol > 0 ? 1.0 - (4.0 / 3.0) * M_PI * excl * excl * excl / (L * L * L) : 1.0;
    int n_side = std::max(1, int(std::ceil(std::cbrt(nmol / std::max(vol_frac, 0.05)))));
    std::vector<int> sites;
    for (;; ++n_side) {
      sites = shuffled_sites(g, n_side, L, excl, cx, cy, cz);
      if ((int)sites.size() >= nmol) break;
    }
    const double a = L / n_side;
    for (int m = 0; m < nmol; ++m) {
      const int s = sites[m];
      const int i = s % n_side, j = (s / n_side) % n_side, k = s / (n_side * n_side);
      const double ox = (i + 0.5) * a + 0.3 * (2.0 * g.uniform() - 1.0);
      const double oy = (j + 0.5) * a + 0.3 * (2.0 * g.uniform() - 1.0);
      const double oz = (k + 0.5) * a + 0.3 * (2.0 * g.uniform() - 1.0);
      double R[3][3];
      random_rotation(g, R);
      const int io = 3 * m;
      x[io] = ox;
      y[io] = oy;
      z[io] = oz;
      double hv[2][3];
      for (int h = 0; h < 2; ++h) {
        for (int c = 0; c < 3; ++c)
          hv[h][c] = R[c][0] * hloc[h][0] + R[c][1] * hloc[h][1] + R[c][2] * hloc[h][2];
        x[io + 1 + h] = ox + hv[h][0];
        y[io + 1 + h] = oy + hv[h][1];
        z[io + 1 + h] = oz + hv[h][2];
      }
      const double bis[3] = {R[0][2], R[1][2], R[2][2]};
      for (int t = 0; t < 3; ++t) {
        double* rp = rpole + (size_t)(io + t) * 13;
        rp[0] = t == 0 ? kQO : kQH;
        for (int c = 0; c < 3; ++c) rp[1 + c] = t == 0 ? kDipO * bis[c] : kDipH * hv[t - 1][c] / kBond;
        random_quadrupole(g, rp);
        polarity[io + t] = t == 0 ? kPolO : kPolH;
        thole[io + t] = kThole;
      }
      const int pr[3][2] = {{io, io + 1}, {io, io + 2}, {io + 1, io + 2}};
      for (int p = 0; p < 3; ++p) {
        scaled_ik[2 * nsc] = pr[p][0];
        scaled_ik[2 * nsc + 1] = pr[p][1];
        scaled_dpu[3 * nsc] = 0.0;
        scaled_dpu[3 * nsc + 1] = 0.0;
        scaled_dpu[3 * nsc + 2] = 1.0;
        ++nsc;
      }
    }
  }
  // ---- solute blob: jittered lattice points inside the sphere
  if (nsol > 0) {
    const double dens = nsol / ((4.0 / 3.0) * M_PI * blob_radius * blob_radius * blob_radius);
    double a = std::cbrt(1.0 / dens);
    std::vector<double> px, py, pz;
    for (;; a *= 0.97) {
      px.clear();
      py.clear();
      pz.clear();
      const int m = int(std::ceil(blob_radius / a)) + 1;
      for (int k = -m; k <= m; ++k)
        for (int j = -m; j <= m; ++j)
          for (int i = -m; i <= m; ++i) {
            const double dx = i * a, dy = j * a, dz = k * a;
            if (dx * dx + dy * dy + dz * dz <= blob_radius * blob_radius) {
              px.push_back(cx + dx);
              py.push_back(cy + dy);
              pz.push_back(cz + dz);
            }
          }
      if ((int)px.size() >= nsol) break;
    }
    std::vector<int> order(px.size());
    for (size_t q = 0; q < order.size(); ++q) order[q] = (int)q;
    for (size_t q = order.size(); q > 1; --q) std::swap(order[q - 1], order[g.below(q)]);
    const double jit = std::min(0.3, 0.25 * a);
    double qsum = 0.0;
    const int base = 3 * nmol;
    for (int s = 0; s < nsol; ++s) {
      const int q = order[s];
      const int ia = base + s;
      x[ia] = px[q] + jit * (2.0 * g.uniform() - 1.0);
      y[ia] = py[q] + jit * (2.0 * g.uniform() - 1.0);
      z[ia] = pz[q] + jit * (2.0 * g.uniform() - 1.0);
      double* rp = rpole + (size_t)ia * 13;
      rp[0] = g.normal(0.3);
      qsum += rp[0];
      for (int c = 0; c < 3; ++c) rp[1 + c] = g.normal(0.1);
      random_quadrupole(g, rp);
      polarity[ia] = 1.0 + 0.75 * g.uniform();
      thole[ia] = kThole;
    }
    const double mean = qsum / nsol;
    for (int s = 0; s < nsol; ++s) rpole[(size_t)(base + s) * 13] -= mean;
  }
  *nscaled = nsc;
  return THP_OK;
}
