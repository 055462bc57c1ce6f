// Deterministic synthetic volumes (host).  The image is z-fastest:
// img[(x*Y + y)*Z + z], the layout the reference hands to its extractors
// (types/src/Isosurface.cpp:185-200).  Values are computed in double and
// rounded to float once, so every machine running this library produces
// the same bytes for the same seed.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "iso_amd.h"

namespace {

// SplitMix64 (Steele, Lea, Flood 2014).
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1p-53; }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// counter-based uniform in [0,1) for (seed, index)
double hash_uniform(uint64_t seed, uint64_t idx) {
  return (double)(mix64(seed * 0x9E3779B97F4A7C15ull + idx + 1) >> 11) * 0x1p-53;
}

template <class Fn>
void parallel_x(int X, Fn fn) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if ((int)nt > X) nt = (unsigned)X;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([=]() {
      for (int x = (int)t; x < X; x += (int)nt) fn(x);
    });
  for (auto &t : th) t.join();
}

// value noise on an integer lattice of spacing `cell`, smoothstep-weighted
double value_noise(uint64_t seed, double x, double y, double z, double cell) {
  double fx = x / cell, fy = y / cell, fz = z / cell;
  double ix = std::floor(fx), iy = std::floor(fy), iz = std::floor(fz);
  double tx = fx - ix, ty = fy - iy, tz = fz - iz;
  tx = tx * tx * (3 - 2 * tx);
  ty = ty * ty * (3 - 2 * ty);
  tz = tz * tz * (3 - 2 * tz);
  auto lat = [&](int64_t a, int64_t b, int64_t c) {
    uint64_t key = ((uint64_t)(a & 0x1fffff) << 42) | ((uint64_t)(b & 0x1fffff) << 21) |
                   (uint64_t)(c & 0x1fffff);
    return 2.0 * hash_uniform(seed, key) - 1.0;
  };
  int64_t a = (int64_t)ix, b = (int64_t)iy, c = (int64_t)iz;
  double v000 = lat(a, b, c), v100 = lat(a + 1, b, c), v010 = lat(a, b + 1, c),
         v110 = lat(a + 1, b + 1, c), v001 = lat(a, b, c + 1), v101 = lat(a + 1, b, c + 1),
         v011 = lat(a, b + 1, c + 1), v111 = lat(a + 1, b + 1, c + 1);
  double x00 = v000 + tx * (v100 - v000), x10 = v010 + tx * (v110 - v010);
  double x01 = v001 + tx * (v101 - v001), x11 = v011 + tx * (v111 - v011);
  double y0 = x00 + ty * (x10 - x00), y1 = x01 + ty * (x11 - x01);
  return y0 + tz * (y1 - y0);
}

}  // namespace

extern "C" {

int iso_gen_blobs(float *img, int X, int Y, int Z, int nblobs, float noise, uint64_t seed) {
  if (!img || X <= 0 || Y <= 0 || Z <= 0 || nblobs < 0) return ISO_ERR_INVALID_ARG;
  SplitMix64 rng(seed);
  std::vector<double> cx(nblobs), cy(nblobs), cz(nblobs), inv2s2(nblobs), amp(nblobs);
  for (int b = 0; b < nblobs; ++b) {
    cx[b] = rng.uniform(0, X);
    cy[b] = rng.uniform(0, Y);
    cz[b] = rng.uniform(0, Z);
    double s = rng.uniform(4, 12);
    inv2s2[b] = 1.0 / (2 * s * s);
    amp[b] = rng.uniform(0.5, 1.5);
  }
  uint64_t nseed = rng.next();
  parallel_x(X, [&](int x) {
    for (int y = 0; y < Y; ++y)
      for (int z = 0; z < Z; ++z) {
        double v = 0;
        for (int b = 0; b < nblobs; ++b) {
          double dx = x - cx[b], dy = y - cy[b], dz = z - cz[b];
          v += amp[b] * std::exp(-(dx * dx + dy * dy + dz * dz) * inv2s2[b]);
        }
        uint64_t idx = ((uint64_t)x * Y + y) * Z + z;
        v += noise * (2.0 * hash_uniform(nseed, idx) - 1.0);
        img[idx] = (float)v;
      }
  });
  return ISO_OK;
}

int iso_gen_front(float *img, int X, int Y, int Z, uint64_t seed) {
  if (!img || X <= 0 || Y <= 0 || Z <= 0) return ISO_ERR_INVALID_ARG;
  const double w = 8.0, A = 20.0, ly = Y / 2.0, lz = Z / 2.0;
  const double pi = 3.14159265358979323846;
  SplitMix64 rng(seed);
  uint64_t s0 = rng.next(), s1 = rng.next(), s2 = rng.next();
  parallel_x(X, [&](int x) {
    for (int y = 0; y < Y; ++y) {
      double x0y = A * std::sin(2 * pi * y / ly);
      for (int z = 0; z < Z; ++z) {
        double x0 = X / 2.0 + x0y * std::sin(2 * pi * z / lz);
        double n = (value_noise(s0, x, y, z, 32.0) + 0.5 * value_noise(s1, x, y, z, 16.0) +
                    0.25 * value_noise(s2, x, y, z, 8.0)) /
                   1.75;
        double v = 0.5 + 0.5 * std::tanh((x - x0) / w) + 0.1 * n;
        v = v < 0 ? 0 : (v > 1 ? 1 : v);
        img[((uint64_t)x * Y + y) * Z + z] = (float)v;
      }
    }
  });
  return ISO_OK;
}

int iso_gen_sines(float *img, int X, int Y, int Z, uint64_t seed) {
  if (!img || X <= 0 || Y <= 0 || Z <= 0) return ISO_ERR_INVALID_ARG;
  SplitMix64 rng(seed);
  double a[4], b[4], c[4];
  for (int m = 0; m < 4; ++m) {
    a[m] = rng.uniform(0.05, 0.4);
    b[m] = rng.uniform(0.05, 0.4);
    c[m] = rng.uniform(0.05, 0.4);
  }
  parallel_x(X, [&](int x) {
    for (int y = 0; y < Y; ++y)
      for (int z = 0; z < Z; ++z) {
        double v = 0;
        for (int m = 0; m < 4; ++m) v += std::sin(a[m] * x) * std::sin(b[m] * y) * std::sin(c[m] * z);
        img[((uint64_t)x * Y + y) * Z + z] = (float)v;
      }
  });
  return ISO_OK;
}

}  // extern "C"
