// Representative isovalues from a similarity map (host), and the multivariate
// time-varying workload generator.
//
// Representative selection restates SimilarityMap::findRepresentativeIsovalues
// (types/src/SimilarityMap.cpp:88-227, Bruckner & Moeller 2010):
//   prioritize(S): pick the element of S with the largest mean similarity to S
//     (first maximum in S's order), give it priority |S| * mean, then recurse
//     on the elements with a smaller map index and on those with a larger one;
//   select: cnt times, take the highest priority, divide every remaining
//     priority by (1 + similarity to the taken element).
// The reference keeps the candidates in a Boost Fibonacci heap; equal
// priorities are taken here in ascending map index (its tie order is not
// pinned by anything in the reference).  The returned ids are sorted
// ascending, the order in which the reference names its distancefield-<n>.df
// outputs (runtime/src/Single.cpp:476-545, INPUT_ORDER).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "iso_amd.h"

namespace {

struct Cand {
  float priority;
  int id;
  int map_id;
};

void prioritize(const float *map, int m, std::vector<int> elems, std::vector<Cand> &out,
                const int *ids) {
  if (elems.empty()) return;
  const int n = (int)elems.size();
  int best = -1;
  float best_avg = -1.0f;
  for (int a = 0; a < n; ++a) {
    float sum = 0.0f;  // float accumulation in set order, as the reference
    for (int b = 0; b < n; ++b) sum += map[(int64_t)elems[a] * m + elems[b]];
    float avg = sum / (float)n;
    if (avg > best_avg) { best_avg = avg; best = a; }
  }
  const int index = elems[best];
  out.push_back(Cand{(float)n * best_avg, ids ? ids[index] : index, index});
  std::vector<int> smaller, bigger;
  for (int e : elems) {
    if (e < index) smaller.push_back(e);
    else if (e != index) bigger.push_back(e);
  }
  prioritize(map, m, smaller, out, ids);
  prioritize(map, m, bigger, out, ids);
}

// cnt rounds of: take the highest priority (ties: lowest map index), divide
// every remaining priority by (1 + its similarity to the taken one).  Fills
// `picked` in selection order.
int select_core(const float *h_map, int m, const int *h_ids, int cnt, std::vector<Cand> &picked) {
  if (!h_map || m <= 0 || cnt <= 0) return ISO_ERR_INVALID_ARG;
  if (cnt > m) cnt = m;
  std::vector<int> all(m);
  for (int i = 0; i < m; ++i) all[i] = i;
  std::vector<Cand> pq;
  prioritize(h_map, m, all, pq, h_ids);
  picked.clear();
  for (int r = 0; r < cnt; ++r) {
    int top = 0;
    for (int i = 1; i < (int)pq.size(); ++i)
      if (pq[i].priority > pq[top].priority ||
          (pq[i].priority == pq[top].priority && pq[i].map_id < pq[top].map_id))
        top = i;
    const Cand c = pq[top];
    picked.push_back(c);
    pq.erase(pq.begin() + top);
    for (Cand &o : pq) o.priority /= (1.0f + h_map[(int64_t)c.map_id * m + o.map_id]);
  }
  return ISO_OK;
}

}  // namespace

extern "C" int iso_select_representatives(const float *h_map, int m, const int *h_ids, int cnt,
                                          int *h_out_ids, float *h_out_priority, int *h_count) {
  if (!h_out_ids || !h_count) return ISO_ERR_INVALID_ARG;
  std::vector<Cand> picked;
  const int st = select_core(h_map, m, h_ids, cnt, picked);
  if (st != ISO_OK) return st;
  std::sort(picked.begin(), picked.end(), [](const Cand &a, const Cand &b) { return a.id < b.id; });
  for (size_t r = 0; r < picked.size(); ++r) {
    h_out_ids[r] = picked[r].id;
    if (h_out_priority) h_out_priority[r] = picked[r].priority;
  }
  *h_count = (int)picked.size();
  return ISO_OK;
}

extern "C" int iso_select_representatives_ordered(const float *h_map, int m, const int *h_ids,
                                                  int cnt, int *h_sel_ids, int *h_sel_map_ids,
                                                  float *h_sel_priority, int *h_count) {
  if (!h_sel_ids || !h_count) return ISO_ERR_INVALID_ARG;
  std::vector<Cand> picked;
  const int st = select_core(h_map, m, h_ids, cnt, picked);
  if (st != ISO_OK) return st;
  for (size_t r = 0; r < picked.size(); ++r) {
    h_sel_ids[r] = picked[r].id;
    if (h_sel_map_ids) h_sel_map_ids[r] = picked[r].map_id;
    if (h_sel_priority) h_sel_priority[r] = picked[r].priority;
  }
  *h_count = (int)picked.size();
  return ISO_OK;
}

// ---- multivariate time-varying volumes ----------------------------------
// Gaussian blobs advected with constant velocities through a periodic box;
// var 0: f, var 1: exp(f), var 2: |grad f|^2 (analytic).  z-fastest image.
namespace {

uint64_t sm64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uni(uint64_t &s, double a, double b) { return a + (b - a) * ((sm64(s) >> 11) * 0x1p-53); }

double wrap(double d, double L) {
  d = std::fmod(d, L);
  if (d > 0.5 * L) d -= L;
  if (d < -0.5 * L) d += L;
  return d;
}

}  // namespace

extern "C" int iso_gen_timeseries(float *img, int X, int Y, int Z, int var, int step, int nblobs,
                                  uint64_t seed) {
  if (!img || X <= 0 || Y <= 0 || Z <= 0 || var < 0 || var > 2 || nblobs <= 0)
    return ISO_ERR_INVALID_ARG;
  uint64_t s = seed;
  std::vector<double> c0(3 * nblobs), v(3 * nblobs), sig(nblobs), amp(nblobs);
  const double L[3] = {(double)X, (double)Y, (double)Z};
  for (int b = 0; b < nblobs; ++b) {
    for (int a = 0; a < 3; ++a) c0[3 * b + a] = uni(s, 0, L[a]);
    for (int a = 0; a < 3; ++a) v[3 * b + a] = uni(s, -2.0, 2.0);  // voxels per step
    sig[b] = uni(s, 6, 16);
    amp[b] = uni(s, 0.5, 1.5);
  }
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (int x = (int)t; x < X; x += (int)nt)
        for (int y = 0; y < Y; ++y)
          for (int z = 0; z < Z; ++z) {
            double f = 0, g[3] = {0, 0, 0};
            const double p[3] = {(double)x, (double)y, (double)z};
            for (int b = 0; b < nblobs; ++b) {
              double d[3], r2 = 0;
              for (int a = 0; a < 3; ++a) {
                d[a] = wrap(p[a] - (c0[3 * b + a] + v[3 * b + a] * step), L[a]);
                r2 += d[a] * d[a];
              }
              const double s2 = sig[b] * sig[b];
              const double e = amp[b] * std::exp(-r2 / (2 * s2));
              f += e;
              for (int a = 0; a < 3; ++a) g[a] -= e * d[a] / s2;
            }
            double out = var == 0 ? f : (var == 1 ? std::exp(f) : g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            img[((int64_t)x * Y + y) * Z + z] = (float)out;
          }
    });
  for (auto &t : th) t.join();
  return ISO_OK;
}
