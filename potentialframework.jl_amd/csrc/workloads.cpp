// Host-only generator of the fitting-set workload: ncfg periodic
// configurations with log-uniform atom counts, randomly sheared cubic cells at
// a given density, uniform placement with rejection below rmin (random
// sequential addition accelerated by a fractional-coordinate bin grid).
// Deterministic for a given seed (xoshiro256** seeded through splitmix64).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (auto& v : s) {
      seed += 0x9e3779b97f4a7c15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      v = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

void inv3(const double* A, double* B) {  // column-major
  const double a = A[0], b = A[3], c = A[6], d = A[1], e = A[4], f = A[7], g = A[2], h = A[5], i = A[8];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  const double m[9] = {(e * i - f * h), -(d * i - f * g), (d * h - e * g), -(b * i - c * h), (a * i - c * g),
                       -(a * h - b * g), (b * f - c * e), -(a * f - c * d), (a * e - b * d)};
  for (int k = 0; k < 9; ++k) B[k] = m[k] / det;
}

// One configuration; returns attempts or -1 when saturated.
int64_t place(Rng& rng, int n, double density, double rmin, double shear, int64_t max_attempts, double* X,
              double* H) {
  const double L = std::cbrt(n / density);
  for (int bcol = 0; bcol < 3; ++bcol)
    for (int arow = 0; arow < 3; ++arow) {
      const double sab = arow == bcol ? 0.0 : -shear + 2.0 * shear * rng.uniform();
      H[arow + 3 * bcol] = L * ((arow == bcol ? 1.0 : 0.0) + sab);
    }
  double Hi[9];
  inv3(H, Hi);
  int nb[3];
  for (int a = 0; a < 3; ++a) {
    const double reach = rmin * std::sqrt(Hi[a] * Hi[a] + Hi[a + 3] * Hi[a + 3] + Hi[a + 6] * Hi[a + 6]);
    nb[a] = std::max(1, (int)std::floor(1.0 / reach));
  }
  const int nbins = nb[0] * nb[1] * nb[2];
  std::vector<std::vector<int>> bins(nbins);
  std::vector<double> F(3 * (size_t)n);
  const double r2min = rmin * rmin;
  int64_t attempts = 0;
  for (int placed = 0; placed < n;) {
    if (++attempts > max_attempts) return -1;
    double f[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
    int key[3];
    for (int a = 0; a < 3; ++a) key[a] = std::min(nb[a] - 1, (int)(f[a] * nb[a]));
    bool ok = true;
    for (int dz = -1; dz <= 1 && ok; ++dz)
      for (int dy = -1; dy <= 1 && ok; ++dy)
        for (int dx = -1; dx <= 1 && ok; ++dx) {
          const int o[3] = {dx, dy, dz};
          bool dup = false;  // with fewer than 3 bins an offset revisits a bin
          for (int a = 0; a < 3; ++a)
            if (nb[a] < 3 && o[a] != 0 && (nb[a] == 1 || o[a] == 1)) dup = true;
          if (dup) continue;
          int b[3];
          for (int a = 0; a < 3; ++a) b[a] = ((key[a] + o[a]) % nb[a] + nb[a]) % nb[a];
          for (int j : bins[b[0] + nb[0] * (b[1] + nb[1] * b[2])]) {
            double df[3];
            for (int a = 0; a < 3; ++a) {
              df[a] = F[3 * j + a] - f[a];
              df[a] -= std::nearbyint(df[a]);
            }
            double r2 = 0;
            for (int a = 0; a < 3; ++a) {
              const double d = H[a] * df[0] + H[a + 3] * df[1] + H[a + 6] * df[2];
              r2 += d * d;
            }
            if (r2 < r2min) {
              ok = false;
              break;
            }
          }
        }
    if (!ok) continue;
    for (int a = 0; a < 3; ++a) F[3 * placed + a] = f[a];
    bins[key[0] + nb[0] * (key[1] + nb[1] * key[2])].push_back(placed);
    for (int a = 0; a < 3; ++a) X[3 * placed + a] = H[a] * f[0] + H[a + 3] * f[1] + H[a + 6] * f[2];
    ++placed;
  }
  return attempts;
}

}  // namespace

extern "C" {

// Pass X == nullptr to receive only the atom offsets (ncfg + 1). Returns 0,
// 1 for invalid arguments, 2 when a configuration saturates (more than
// max_attempts_per_atom * n attempts).
int pfw_wl_fitting_set(int ncfg, int nmin, int nmax, double density, double rmin, double shear, uint64_t seed,
                       int64_t max_attempts_per_atom, int64_t* offsets, double* X, int32_t* species, double* H,
                       int64_t* attempts) {
  if (ncfg <= 0 || nmin <= 0 || nmax < nmin || !(density > 0) || !(rmin >= 0) || !offsets) return 1;
  if (rmin > 0.45 * std::cbrt(nmin / density)) return 1;  // the nearest-image rounding would be inexact
  Rng sizes(seed);
  offsets[0] = 0;
  for (int c = 0; c < ncfg; ++c) {
    const double u = std::log((double)nmin) + (std::log((double)nmax) - std::log((double)nmin)) * sizes.uniform();
    offsets[c + 1] = offsets[c] + (int64_t)std::floor(std::exp(u));
  }
  if (!X) return 0;
  int64_t total = 0;
  for (int c = 0; c < ncfg; ++c) {
    Rng rng(seed * 1000003ull + 7919ull * (uint64_t)(c + 1));
    const int n = (int)(offsets[c + 1] - offsets[c]);
    const int64_t a = place(rng, n, density, rmin, shear, max_attempts_per_atom * n, X + 3 * offsets[c], H + 9 * c);
    if (a < 0) return 2;
    total += a;
    for (int i = 0; i < n; ++i) species[offsets[c] + i] = (int32_t)(rng.next() >> 63);
  }
  if (attempts) *attempts = total;
  return 0;
}

}  // extern "C"
