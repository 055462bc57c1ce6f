/*
 * CPU oracle for the ISM hot path -- TEST INFRASTRUCTURE ONLY (see iso_oracle.h).
 * Reference paths are relative to isosurfaces_lib/ of mimre25/isosurface-analysis.
 */
#include "iso_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Corner numbering and edge end points of the reference's cell
 * (isosurfaces_cuda/src/MarchingCube.cu:208-265). */
static const int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                 {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
static const int EDGE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
/* The six faces as corner cycles. */
static const int FACE[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                               {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};

void orc_ingest(const float *file, int X, int Y, int Z, float cfg_min, float cfg_max,
                float *img, float *real_min, float *real_max) {
  float mx = -100000, mn = 100000;
  for (int x = 0; x < X; ++x)
    for (int y = 0; y < Y; ++y)
      for (int z = 0; z < Z; ++z) {
        float v = file[(int64_t)x + (int64_t)y * X + (int64_t)z * X * Y];
        img[((int64_t)x * Y + y) * Z + z] = v;
        if (v > mx && v <= cfg_max) mx = v;
        if (v < mn && v >= cfg_min) mn = v;
      }
  *real_min = mn;
  *real_max = mx;
}

void orc_equalized_isovalues(const float *data, int64_t n, float gmin, float gmax, int k,
                             float *out) {
  const int B = 65536;
  int *hist = (int *)calloc(B, sizeof(int));
  float oneovergrange = 1.0 / (gmax - gmin);
  float dataSize = (float)n;
  for (int64_t i = 0; (float)i < dataSize; i++) {
    int index = (int)(B * (data[i] - gmin) * oneovergrange);
    if (index > B - 1) index = B - 1;
    if (index < 0) index = 0;
    hist[index] += 1;
  }
  double sum = 0.0, acc, threshold;
  for (int i = 0; i < B; i++) sum += hist[i];
  acc = 0;
  int index = 0;
  threshold = sum / (double)k;
  for (int i = 0; i < k; i++) {
    while (((acc == 0.0) || ((acc + hist[index]) < threshold)) && (index < B - 1)) {
      acc += hist[index];
      hist[index] = i;
      index++;
    }
    sum -= acc;
    acc = 0.0;
    if (i < k - 1) threshold = sum / (double)(k - (i + 1));
  }
  int idx = 0, lastIdx = 0;
  float step = (gmax - gmin) / B;
  for (int i = 0; i < k; ++i) {
    while (idx < B && hist[idx] <= i) ++idx;
    int range = idx - lastIdx;
    int mid = lastIdx + range / 2;
    out[i] = gmin + step * mid;
    lastIdx = idx;
  }
  free(hist);
}

static int edge_between(int a, int b) {
  for (int e = 0; e < 12; ++e)
    if ((EDGE[e][0] == a && EDGE[e][1] == b) || (EDGE[e][0] == b && EDGE[e][1] == a))
      return e;
  return -1;
}

static int uf_find(int *p, int x) {
  while (p[x] != x) x = p[x] = p[p[x]];
  return x;
}

void orc_mc_tables(uint16_t edge_mask[256], uint8_t num_verts[256]) {
  for (int c = 0; c < 256; ++c) {
    int crossed = 0, ne = 0;
    for (int e = 0; e < 12; ++e)
      if (((c >> EDGE[e][0]) & 1) != ((c >> EDGE[e][1]) & 1)) { crossed |= 1 << e; ++ne; }
    int parent[12];
    for (int e = 0; e < 12; ++e) parent[e] = e;
    for (int f = 0; f < 6; ++f) {
      int fe[4], nc = 0, ce[4];
      for (int i = 0; i < 4; ++i) {
        fe[i] = edge_between(FACE[f][i], FACE[f][(i + 1) & 3]);
        if (crossed >> fe[i] & 1) ce[nc++] = fe[i];
      }
      if (nc == 2) {
        parent[uf_find(parent, ce[0])] = uf_find(parent, ce[1]);
      } else if (nc == 4) {
        /* ambiguous face: cut off every corner whose bit is set */
        for (int i = 0; i < 4; ++i)
          if ((c >> FACE[f][i]) & 1) {
            int e1 = fe[(i + 3) & 3], e2 = fe[i];
            parent[uf_find(parent, e1)] = uf_find(parent, e2);
          }
      }
    }
    int polys = 0;
    for (int e = 0; e < 12; ++e)
      if ((crossed >> e & 1) && uf_find(parent, e) == e) ++polys;
    edge_mask[c] = (uint16_t)crossed;
    num_verts[c] = (uint8_t)(3 * (ne - 2 * polys));
  }
}

int64_t orc_mc_ncls_compat(int gx, int gy, int gz) {
  /* MarchingCube.cpp:599-608: int numVoxels; grid(numVoxels/threads) */
  int64_t n = (int64_t)gx * gy * gz;
  int64_t bx = n / 128, by = 1;
  if (bx > 65535) { by = bx / 32768; bx = 32768; }
  int64_t c = 128 * bx * by;
  return c < n ? c : n;
}

typedef struct { float x, y, z; } f3;

static int f3_cmp(const void *a, const void *b) {
  const f3 *p = (const f3 *)a, *q = (const f3 *)b;
  if (p->x != q->x) return p->x < q->x ? -1 : 1;
  if (p->y != q->y) return p->y < q->y ? -1 : 1;
  if (p->z != q->z) return p->z < q->z ? -1 : 1;
  return 0;
}

int64_t orc_mc(const float *vol, int gx, int gy, int gz, float iso, int64_t ncls,
               int interior, float *out_xyz, int64_t cap, int64_t stats[2]) {
  uint16_t emask[256];
  uint8_t nverts[256];
  orc_mc_tables(emask, nverts);
  int64_t n = (int64_t)gx * gy * gz;
  if (ncls > n) ncls = n;
  int64_t vcap = 1 << 16, nv = 0, active = 0, total = 0;
  f3 *v = (f3 *)malloc(sizeof(f3) * vcap);
  if (!v) return -1;
  for (int64_t i = 0; i < ncls; ++i) {
    int64_t p[3] = {i % gx, (i / gx) % gy, i / ((int64_t)gx * gy)};
    int64_t dims[3] = {gx, gy, gz};
    if (interior && (p[0] >= gx - 1 || p[1] >= gy - 1 || p[2] >= gz - 1)) continue;
    float f[8];
    for (int k = 0; k < 8; ++k) {
      int64_t q[3];
      for (int a = 0; a < 3; ++a) {
        q[a] = p[a] + CORNER[k][a];
        if (q[a] > dims[a] - 1) q[a] = dims[a] - 1; /* sampleVolume clamp (Q6) */
      }
      f[k] = vol[q[2] * gx * gy + q[1] * gx + q[0]];
    }
    int c = 0;
    for (int k = 0; k < 8; ++k) c |= (f[k] < iso) << k;
    total += nverts[c];
    if (nverts[c] == 0) continue;
    ++active;
    for (int e = 0; e < 12; ++e) {
      if (!(emask[c] >> e & 1)) continue;
      int a = EDGE[e][0], b = EDGE[e][1];
      float t = (iso - f[a]) / (f[b] - f[a]); /* vertexInterp, MarchingCube.cu:159-164 */
      float out[3];
      for (int ax = 0; ax < 3; ++ax) {
        float pa = (float)p[ax] + (float)CORNER[a][ax];
        float pb = (float)p[ax] + (float)CORNER[b][ax];
        out[ax] = pa + t * (pb - pa); /* lerp(a,b,t) = a + t*(b-a) */
      }
      if (nv == vcap) {
        vcap *= 2;
        f3 *nvp = (f3 *)realloc(v, sizeof(f3) * vcap);
        if (!nvp) { free(v); return -1; }
        v = nvp;
      }
      v[nv].x = out[0]; v[nv].y = out[1]; v[nv].z = out[2];
      ++nv;
    }
  }
  qsort(v, (size_t)nv, sizeof(f3), f3_cmp);
  int64_t u = 0;
  for (int64_t i = 0; i < nv; ++i) {
    if (u > 0 && f3_cmp(&v[i], &v[u - 1]) == 0) continue;
    v[u++] = v[i];
  }
  for (int64_t i = 0; i < u && i < cap; ++i) {
    out_xyz[3 * i] = v[i].x; out_xyz[3 * i + 1] = v[i].y; out_xyz[3 * i + 2] = v[i].z;
  }
  free(v);
  if (stats) { stats[0] = active; stats[1] = total; }
  return u;
}

int64_t orc_approx(const float *img, int X, int Y, int Z, float iso, int32_t *out_xyz,
                   int64_t cap) {
  int64_t m = 0;
  int64_t YZ = (int64_t)Y * Z;
  for (int x = 0; x < X - 1; ++x)
    for (int y = 0; y < Y - 1; ++y)
      for (int z = 0; z < Z - 1; ++z) {
        int64_t lu = x * YZ + (int64_t)y * Z + z;
        int64_t id[8] = {lu, lu + 1, lu + Z, lu + Z + 1,
                         lu + YZ, lu + 1 + YZ, lu + Z + YZ, lu + Z + 1 + YZ};
        int above = 0, below = 0;
        for (int k = 0; k < 8; ++k) {
          float v = img[id[k]];
          above |= v > iso;
          below |= v < iso;
        }
        if (above && below) {
          if (m < cap) {
            out_xyz[3 * m] = x; out_xyz[3 * m + 1] = y; out_xyz[3 * m + 2] = z;
          }
          ++m;
        }
      }
  return m;
}

static int dim_max(int X, int Y, int Z) {
  int m = X > Y ? X : Y;
  return m > Z ? m : Z;
}

void orc_normalize_f3(const float *pts, int64_t n, int X, int Y, int Z, float *out) {
  float dm = (float)dim_max(X, Y, Z);
  for (int64_t i = 0; i < 3 * n; ++i) out[i] = pts[i] / dm;
}

void orc_normalize_i3(const int32_t *pts, int64_t n, int X, int Y, int Z, float *out) {
  float dm = (float)dim_max(X, Y, Z);
  for (int64_t i = 0; i < 3 * n; ++i) out[i] = (float)pts[i] / dm;
}

void orc_sample_point(int64_t idx, int X, int Y, int Z, int ds, float out[3]) {
  int Yd = Y / ds, Zd = Z / ds;
  int z = (int)(idx % Zd);
  int y = (int)((idx / Zd) % Yd);
  int x = (int)(idx / ((int64_t)Yd * Zd));
  out[0] = (float)(x * ds + ds / 2) / (float)X;
  out[1] = (float)(y * ds + ds / 2) / (float)Y;
  out[2] = (float)(z * ds + ds / 2) / (float)Z;
}

float orc_d2(const float s[3], const float p[3]) {
  float dx = s[0] - p[0], dy = s[1] - p[1], dz = s[2] - p[2];
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

static float df_one(const float *pts, int64_t n, const float s[3]) {
  float best = INFINITY;
  for (int64_t i = 0; i < n; ++i) {
    float d = orc_d2(s, pts + 3 * i);
    if (d < best) best = d;
  }
  return n == 1 ? best : sqrtf(best);
}

void orc_df_brute(const float *pts, int64_t n, int X, int Y, int Z, int ds, float *out,
                  int nthreads) {
  int64_t F = (int64_t)(X / ds) * (Y / ds) * (Z / ds);
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (int64_t i = 0; i < F; ++i) {
    float s[3];
    orc_sample_point(i, X, Y, Z, ds, s);
    out[i] = df_one(pts, n, s);
  }
}

void orc_df_brute_subset(const float *pts, int64_t n, int X, int Y, int Z, int ds,
                         const int64_t *idx, int64_t ns, float *out, int nthreads) {
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t i = 0; i < ns; ++i) {
    float s[3];
    orc_sample_point(idx[i], X, Y, Z, ds, s);
    out[i] = df_one(pts, n, s);
  }
}

/* ---- CPU hierarchy for the distance-field baseline ---- */
typedef struct {
  float lo[3], hi[3];
  int64_t a, b; /* sorted point range [a, b) */
  int64_t left, right; /* children, -1 for a leaf */
} tnode;

typedef struct {
  const float *p; /* sorted points, xyz */
  tnode *nodes;
  int64_t count;
} tree;

static uint32_t spread10(float c) {
  c = fminf(fmaxf(c * 1024.0f, 0.0f), 1023.0f);
  uint32_t v = (uint32_t)c;
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

typedef struct { uint32_t code; int64_t i; } coded;
static int coded_cmp(const void *a, const void *b) {
  const coded *x = (const coded *)a, *y = (const coded *)b;
  if (x->code != y->code) return x->code < y->code ? -1 : 1;
  return x->i < y->i ? -1 : (x->i > y->i);
}

static int64_t tree_build(tree *t, int64_t a, int64_t b) {
  int64_t id = t->count++;
  tnode *nd = &t->nodes[id];
  nd->a = a; nd->b = b;
  for (int k = 0; k < 3; ++k) { nd->lo[k] = INFINITY; nd->hi[k] = -INFINITY; }
  for (int64_t i = a; i < b; ++i)
    for (int k = 0; k < 3; ++k) {
      float v = t->p[3 * i + k];
      if (v < nd->lo[k]) nd->lo[k] = v;
      if (v > nd->hi[k]) nd->hi[k] = v;
    }
  if (b - a <= 8) { nd->left = nd->right = -1; return id; }
  int64_t m = a + (b - a) / 2;
  int64_t l = tree_build(t, a, m);
  int64_t r = tree_build(t, m, b);
  t->nodes[id].left = l;
  t->nodes[id].right = r;
  return id;
}

static float box_d2(const float s[3], const tnode *n) {
  float d[3];
  for (int k = 0; k < 3; ++k) {
    float g = fmaxf(fmaxf(n->lo[k] - s[k], s[k] - n->hi[k]), 0.0f);
    d[k] = g;
  }
  return fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
}

static void tree_query(const tree *t, int64_t id, const float s[3], float *best) {
  const tnode *n = &t->nodes[id];
  if (n->left < 0) {
    for (int64_t i = n->a; i < n->b; ++i) {
      float d = orc_d2(s, t->p + 3 * i);
      if (d < *best) *best = d;
    }
    return;
  }
  float dl = box_d2(s, &t->nodes[n->left]), dr = box_d2(s, &t->nodes[n->right]);
  int64_t first = dl <= dr ? n->left : n->right, second = dl <= dr ? n->right : n->left;
  float d1 = dl <= dr ? dl : dr, d2 = dl <= dr ? dr : dl;
  if (d1 < *best) tree_query(t, first, s, best);
  if (d2 < *best) tree_query(t, second, s, best);
}

static void df_tree_seeded(const float *pts, int64_t n, int X, int Y, int Z, int ds,
                           const int64_t *sample_idx, int64_t ns, float *out, int nthreads,
                           int presample);

void orc_df_tree(const float *pts, int64_t n, int X, int Y, int Z, int ds,
                 const int64_t *sample_idx, int64_t ns, float *out, int nthreads) {
  df_tree_seeded(pts, n, X, Y, Z, ds, sample_idx, ns, out, nthreads, 0);
}

void orc_df_tree_ref(const float *pts, int64_t n, int X, int Y, int Z, int ds,
                     const int64_t *sample_idx, int64_t ns, float *out, int nthreads) {
  df_tree_seeded(pts, n, X, Y, Z, ds, sample_idx, ns, out, nthreads, 1500);
}

static void df_tree_seeded(const float *pts, int64_t n, int X, int Y, int Z, int ds,
                           const int64_t *sample_idx, int64_t ns, float *out, int nthreads,
                           int presample) {
  coded *c = (coded *)malloc(sizeof(coded) * (size_t)n);
  float *sp = (float *)malloc(sizeof(float) * 3 * (size_t)n);
  tree t;
  t.nodes = (tnode *)malloc(sizeof(tnode) * (size_t)(2 * n + 1));
  t.count = 0;
  for (int64_t i = 0; i < n; ++i) {
    c[i].code = spread10(pts[3 * i]) * 4 + spread10(pts[3 * i + 1]) * 2 + spread10(pts[3 * i + 2]);
    c[i].i = i;
  }
  qsort(c, (size_t)n, sizeof(coded), coded_cmp);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) sp[3 * i + k] = pts[3 * c[i].i + k];
  t.p = sp;
  tree_build(&t, 0, n);
  int64_t F = (int64_t)(X / ds) * (Y / ds) * (Z / ds);
  int64_t m = sample_idx ? ns : F;
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (int64_t i = 0; i < m; ++i) {
    float s[3];
    orc_sample_point(sample_idx ? sample_idx[i] : i, X, Y, Z, ds, s);
    float best = INFINITY;
    if (presample > 0) {
      /* the reference's upper bound from S = min(presample, P) of the sorted
       * points, ids (uint32)(i * P) / S in 32-bit arithmetic (BVH.cu:640-644) */
      const uint32_t S = (uint32_t)(n < presample ? n : presample);
      for (uint32_t q = 0; q < S; ++q) {
        const uint32_t id = (uint32_t)(q * (uint32_t)n) / S;
        const float d2 = orc_d2(s, sp + 3 * (int64_t)id);
        if (d2 < best) best = d2;
      }
    }
    tree_query(&t, 0, s, &best);
    out[i] = n == 1 ? best : sqrtf(best);
  }
  free(c);
  free(sp);
  free(t.nodes);
}

void orc_interval(const float *f, int64_t F, float *mn, float *mx) {
  float lo = FLT_MAX, hi = FLT_MIN;
  for (int64_t i = 0; i < F; ++i) {
    if (f[i] < lo) lo = f[i];
    if (f[i] > hi) hi = f[i];
  }
  *mn = lo;
  *mx = hi;
}

static int find_bucket(float v, float mn, float step, int H) {
  if (step == 0) return 0;
  int b = (int)((v - mn) / step);
  return b >= H ? H - 1 : b;
}

void orc_hist_pair(const float *f1, const float *f2, int64_t processed, int H, float mn1,
                   float mx1, float mn2, float mx2, uint32_t *hist) {
  float mn = mn1 < mn2 ? mn1 : mn2;
  float mx = mx1 > mx2 ? mx1 : mx2;
  float step = (mx - mn) / (float)H;
  memset(hist, 0, sizeof(uint32_t) * (size_t)H * H);
  for (int64_t j = 0; j < processed; ++j) {
    int r = find_bucket(f1[j], mn, step, H);
    int c = find_bucket(f2[j], mn, step, H);
    hist[(int64_t)r * H + c]++;
  }
}

int64_t orc_hist_processed(int64_t F) {
  const int64_t C = 1000000;
  if (F > 0 && F % C == 0) return F - C;
  return F;
}

/* logf correctly rounded: the double logarithm rounded once (see
 * orc_logf_cr_margin for why that is the correct rounding for x >= 1). */
static float logf_cr(float x) { return (float)log((double)x); }

/* logf_cr moved by -1, 0 or +1 ulp, a fixed choice per argument and seed:
 * another logf within 1 ulp, as CUDA documents for its own. */
static float logf_ulp(float x, uint64_t seed) {
  float v = logf_cr(x);
  if (seed == 0 || v == 0.0f) return v;
  uint32_t b;
  memcpy(&b, &x, 4);
  uint64_t h = (b ^ seed) * 0x9e3779b97f4a7c15ull;
  h ^= h >> 31;
  h *= 0xbf58476d1ce4e5b9ull;
  h ^= h >> 29;
  int k = (int)(h % 3u);
  return k == 0 ? v : nextafterf(v, k == 1 ? INFINITY : -INFINITY);
}

static float nmi_ref32_seeded(const uint32_t *hist, int H, int64_t N, uint64_t seed) {
  float hX = 0, hY = 0, hXY = 0;
  float numValues = (float)N;
  for (int i = 0; i < H; ++i) {
    uint32_t col = 0, row = 0;
    for (int k = 0; k < H; ++k) { col += hist[(int64_t)k * H + i]; row += hist[(int64_t)i * H + k]; }
    for (int j = 0; j < H; ++j) {
      uint32_t c = hist[(int64_t)i * H + j];
      if (c > 0) { float p = (float)c; hXY = fmaf(-p, logf_ulp(p, seed), hXY); }
    }
    if (col > 0) { float p = (float)col; hX = fmaf(-p, logf_ulp(p, seed), hX); }
    if (row > 0) { float p = (float)row; hY = fmaf(-p, logf_ulp(p, seed), hY); }
  }
  hXY = hXY / numValues + logf_ulp(numValues, seed);
  hX = hX / numValues + logf_ulp(numValues, seed);
  hY = hY / numValues + logf_ulp(numValues, seed);
  float iXY = hX + hY - hXY;
  float val = 2 * iXY / (hX + hY);
  return val != val ? 0.0f : val;
}

float orc_nmi_ref32(const uint32_t *hist, int H, int64_t N) {
  return nmi_ref32_seeded(hist, H, N, 0);
}

float orc_nmi_ref32_ulp(const uint32_t *hist, int H, int64_t N, uint64_t seed) {
  return nmi_ref32_seeded(hist, H, N, seed);
}

int64_t orc_logf_cr_margin(uint32_t lo_bits, uint32_t hi_bits, double rel, int integers_only,
                           uint32_t *first_bad) {
  int64_t bad = 0;
  uint32_t first = 0xffffffffu;
#pragma omp parallel for reduction(+ : bad) schedule(static, 65536)
  for (int64_t u = (int64_t)lo_bits; u <= (int64_t)hi_bits; ++u) {
    uint32_t bits = (uint32_t)u;
    float x;
    memcpy(&x, &bits, 4);
    if (integers_only && x != floorf(x)) continue;
    long double L = logl((long double)x);
    float f = (float)L;
    int ok = 1;
    if (L != 0.0L) {
      float g = nextafterf(f, (long double)f < L ? INFINITY : -INFINITY);
      long double mid = ((long double)f + (long double)g) / 2.0L;
      long double gap = fabsl(L - mid);
      if (gap < (long double)rel * fabsl(L)) ok = 0;
    }
    if (logf_cr(x) != f) ok = 0;
    if (!ok) {
      ++bad;
#pragma omp critical
      if (bits < first) first = bits;
    }
  }
  if (first_bad) *first_bad = first;
  return bad;
}

double orc_nmi_f64(const uint32_t *hist, int H, int64_t N) {
  double sxy = 0, sx = 0, sy = 0;
  for (int i = 0; i < H; ++i) {
    uint64_t row = 0, col = 0;
    for (int k = 0; k < H; ++k) { row += hist[(int64_t)i * H + k]; col += hist[(int64_t)k * H + i]; }
    if (row) sx += (double)row * log((double)row);
    if (col) sy += (double)col * log((double)col);
    for (int j = 0; j < H; ++j) {
      uint32_t c = hist[(int64_t)i * H + j];
      if (c) sxy += (double)c * log((double)c);
    }
  }
  double lnN = log((double)N), dN = (double)N;
  double hx = lnN - sx / dN, hy = lnN - sy / dN, hxy = lnN - sxy / dN;
  double v = 2.0 * (hx + hy - hxy) / (hx + hy);
  return v != v ? 0.0 : v;
}

void orc_ism(const float *fields, int n, int64_t F, int H, int multi, int compat_chunks,
             float *map_ref32, double *map_f64, int nthreads) {
  int nf = multi ? 2 * n : n;
  float *mn = (float *)malloc(sizeof(float) * nf), *mx = (float *)malloc(sizeof(float) * nf);
  for (int i = 0; i < nf; ++i) orc_interval(fields + (int64_t)i * F, F, &mn[i], &mx[i]);
  int64_t processed = compat_chunks ? orc_hist_processed(F) : F;
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    uint32_t *hist = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)H * H);
#pragma omp for schedule(dynamic, 1)
    for (int64_t t = 0; t < (int64_t)n * n; ++t) {
      int i = (int)(t / n), j = (int)(t % n);
      int a = i, b = multi ? j + n : j;
      orc_hist_pair(fields + (int64_t)a * F, fields + (int64_t)b * F, processed, H, mn[a],
                    mx[a], mn[b], mx[b], hist);
      if (map_ref32) map_ref32[(int64_t)i * n + j] = orc_nmi_ref32(hist, H, F);
      if (map_f64) map_f64[(int64_t)i * n + j] = orc_nmi_f64(hist, H, F);
    }
    free(hist);
  }
  free(mn);
  free(mx);
}

void orc_df_voxel_units(const float *pts, int64_t n, int nx, int ny, int nz, int ds, float *out,
                        int nthreads) {
  const int64_t F = (int64_t)nx * ny * nz, run = 2000000;
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
  for (int64_t s = 0; s < F; ++s) {
    const int z = (int)(s % nz), y = (int)((s / nz) % ny), x = (int)(s / ((int64_t)ny * nz));
    const float q[3] = {(float)(x * ds + ds / 2), (float)(y * ds + ds / 2),
                        (float)(z * ds + ds / 2)};
    float m = 1e38f, d = sqrtf(1e38f);
    for (int64_t r0 = 0; r0 < n; r0 += run) {
      if (r0 > 0) m = d * d; /* DistanceField.cu:54-56 */
      for (int64_t i = r0; i < n && i < r0 + run; ++i) {
        float d2 = orc_d2(q, pts + 3 * i);
        if (d2 < m) m = d2;
      }
      d = sqrtf(m);
    }
    out[s] = d;
  }
}

int64_t orc_find_extrema(const float *data, int X, int Y, int Z, float *out, int64_t cap) {
  const int64_t N = (int64_t)X * Y * Z, sy = Y, sz = (int64_t)Y * Z;
  int64_t m = 0;
  for (int64_t idx = 0; idx < N; ++idx) {
    const int z = (int)(idx % Z), y = (int)((idx / Z) % Y), x = (int)(idx / ((int64_t)Y * Z));
    const float c = data[x + y * sy + z * sz];
    int higher = 0, lower = 0, same = 0;
    for (int i = -1; i < 4; ++i)
      for (int j = -1; j < 4; ++j)
        for (int k = -1; k < 4; ++k) {
          const int a = x + i, b = y + j, e = z + k;
          if (a > -1 && b > -1 && e > -1 && a < X && b < Y && e < Z) {
            const float v = data[a + b * sy + e * sz];
            higher += v > c;
            lower += v < c;
            same += v == c;
          }
        }
    float o = c;
    if (same > 5) o = -1e10f;
    if (lower > 0 && higher > 0) o = -1e10f;
    if ((double)o > -1e10) {
      if (m < cap) out[m] = o;
      ++m;
    }
  }
  return m;
}
