/*
 * CPU oracle for the isosurface-similarity-map (ISM) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithm (mimre25/isosurface-analysis, isosurfaces_lib/) used as the checker
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 * The product (libiso_amd.so) never links, loads or calls it.
 *
 * Each function cites the reference file:line it restates.  Paths are relative
 * to isosurfaces_lib/ of the reference.
 *
 * Arithmetic conventions (shared with the HIP kernels, so results are
 * bit-comparable):
 *   - compiled with -ffp-contract=off; the one fused operation is the explicit
 *     squared distance  d2 = fmaf(dz,dz, fmaf(dy,dy, dx*dx)),  dx = s.x - p.x;
 *   - IEEE division and sqrtf (correctly rounded).
 */
#ifndef ISO_ORACLE_H
#define ISO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Volume ingest (filehandler/src/BinaryFloatReader.cpp:29-49,
 * types/src/Isosurface.cpp:32-76,185-200): x-fastest file -> z-fastest image,
 * restricted extremes. */
void orc_ingest(const float *file, int X, int Y, int Z, float cfg_min, float cfg_max,
                float *img, float *real_min, float *real_max);

/* Single::compute_isovalues(uniform=true) + get_uniform_histogram
 * (runtime/src/Single.cpp:286-392), verbatim arithmetic. */
void orc_equalized_isovalues(const float *img, int64_t n, float real_min, float real_max, int k,
                             float *out);

/* Marching-cubes case tables generated from first principles (no table is
 * copied): edge_mask[c] = edges whose end corners differ in case c,
 * num_verts[c] = 3 * (#crossed edges - 2 * #polygons), where polygons are
 * formed by joining crossed edges face by face and, on a face with four
 * crossed edges, cutting off each corner whose bit is set (value < iso).
 * Reference: isosurfaces_cuda/src/tables.h (edgeTable, numVertsTable). */
void orc_mc_tables(uint16_t edge_mask[256], uint8_t num_verts[256]);

/* Number of cells the reference's classify launch covers (Q4):
 * 128 * grid.x * grid.y with grid.x = N/128, split when > 65535.
 * Reference: isosurfaces_cuda/src/MarchingCube.cpp:599-608. */
int64_t orc_mc_ncls_compat(int gx, int gy, int gz);

/* Marching cubes, reference semantics (isosurfaces_cuda/src/MarchingCube.cu:59-331,
 * MarchingCube.cpp:588-710, thrustWrapper.cu:44-56): the buffer is read
 * x-fastest with dims (gx,gy,gz) and clamped corners; cells 0..ncls-1 are
 * classified with f < iso; every crossed edge of an active cell yields
 * lerp(v_a, v_b, (iso-f_a)/(f_b-f_a)) in the cell's own edge orientation;
 * vertices are sorted lexicographically and exact duplicates removed.
 * Writes up to `cap` unique xyz triples, returns the unique count (or -1 on
 * allocation failure).  stats[0] = active cells, stats[1] = total vertices
 * (sum of num_verts over classified cells).  interior != 0 restricts the
 * cells to x<gx-1, y<gy-1, z<gz-1 (no clamped corner): the library's "fixed"
 * mode, which is not a reference behaviour. */
int64_t orc_mc(const float *vol, int gx, int gy, int gz, float iso, int64_t ncls,
               int interior, float *out_xyz, int64_t cap, int64_t stats[2]);

/* Voxel approximation (isosurfaces_cuda/src/IsosurfaceApproximation.cu:67-135,
 * 159-212): cells (x<X-1,y<Y-1,z<Z-1) of the z-fastest image with some corner
 * > iso and some corner < iso, in x-major / z-fastest order. */
int64_t orc_approx(const float *img, int X, int Y, int Z, float iso,
                   int32_t *out_xyz, int64_t cap);

/* Point normalisation p / max(X,Y,Z) (isosurfaces_cuda/funcs.cpp:209-233). */
void orc_normalize_f3(const float *pts, int64_t n, int X, int Y, int Z, float *out);
void orc_normalize_i3(const int32_t *pts, int64_t n, int X, int Y, int Z, float *out);

/* Sample coordinates of DF sample idx (isosurfaces_cuda/src/BVH.cu:393-406). */
void orc_sample_point(int64_t idx, int X, int Y, int Z, int ds, float out[3]);

/* Canonical squared distance between sample s and point p. */
float orc_d2(const float s[3], const float p[3]);

/* Brute-force distance field over the (X/ds,Y/ds,Z/ds) grid, z fastest:
 * sqrtf(min_p d2), or d2 itself when n == 1 (BVH.cu:577-613). The exact
 * minimum the reference's LBVH traversal returns for a well-formed tree. */
void orc_df_brute(const float *pts_norm, int64_t n, int X, int Y, int Z, int ds,
                  float *out, int nthreads);

/* Same, for a subset of sample indices (CPU baseline / spot checks). */
void orc_df_brute_subset(const float *pts_norm, int64_t n, int X, int Y, int Z,
                         int ds, const int64_t *sample_idx, int64_t ns,
                         float *out, int nthreads);

/* Same minimum through a bounding-volume hierarchy, the reference's CPU-side
 * analogue of its LBVH (BVH.cu): points sorted by 30-bit Morton code, a
 * balanced binary tree over the sorted order (leaves of <= 8 points), and a
 * nearest-first recursive search pruned with the canonical point-box bound.
 * sample_idx == NULL means every sample.  Used as the CPU baseline. */
/* The reference's distance-field search order (BVH.cu:577-613, 640-644):
 * the minimum first over a pre-sample of S = min(1500, n) of the Morton-
 * sorted points, ids (uint32)(i * n) / S in 32-bit arithmetic (Q12), then
 * the tree walk pruned by that bound.  Same exact result; the CPU baseline
 * times this one. */
void orc_df_tree_ref(const float *pts_norm, int64_t n, int X, int Y, int Z, int ds,
                     const int64_t *sample_idx, int64_t ns, float *out, int nthreads);

void orc_df_tree(const float *pts_norm, int64_t n, int X, int Y, int Z, int ds,
                 const int64_t *sample_idx, int64_t ns, float *out, int nthreads);

/* calculate_distances_h (isosurfaces_cuda/src/DistanceField.cu:15-98): for
 * each of nx*ny*nz samples (z fastest) at (x*ds + ds/2, ...) in voxel units,
 * the points as given; the minimum starts at 1e38 and the points go in runs
 * of 2,000,000, each run starting from the square of the previous run's
 * sqrtf; out = sqrtf of the last run's minimum. */
void orc_df_voxel_units(const float *pts, int64_t n, int nx, int ny, int nz, int ds, float *out,
                        int nthreads);

/* findExtrema_h (isosurfaces_cuda/src/IsosurfaceExtrema.cu:22-138): voxel idx
 * (z fastest) reads its value and its neighbours' at x + y*Y + z*Y*Z;
 * neighbours x-1..x+3, y-1..y+3, z-1..z+3 inside the grid, itself included;
 * more than 5 equal, or some lower and some higher, give -1e10; the values
 * above -1e10 in voxel order (up to cap written).  Returns their number. */
int64_t orc_find_extrema(const float *data, int X, int Y, int Z, float *out, int64_t cap);

/* DistanceField::getInterval (types/src/DistanceField.cpp:159-200): min
 * starts at FLT_MAX, max at FLT_MIN (Q15). */
void orc_interval(const float *f, int64_t F, float *mn, float *mx);

/* Joint histogram of one pair (isosurfaces_cuda/src/JointHistogram.cu:63-75,
 * 100-162): range [min(mn1,mn2), max(mx1,mx2)], step = (max-min)/H,
 * bin = (int)((v-min)/step) clamped to H-1, step==0 -> 0.
 * hist[row*H+col], row from f1, col from f2.  `processed` voxels are counted
 * (F, or fewer under the 1M-chunk quirk Q16). */
void orc_hist_pair(const float *f1, const float *f2, int64_t processed, int H,
                   float mn1, float mx1, float mn2, float mx2, uint32_t *hist);

/* Voxels the reference's chunk loop actually visits (JointHistogram.cu:372-393):
 * F, minus the final 1,000,000 when F is a positive multiple of 1,000,000. */
int64_t orc_hist_processed(int64_t F);

/* calculateMutualInformation in the reference's float order
 * (JointHistogram.cu:24-60); the reference compiles a - b*c as a fused
 * multiply-add, restated here with fmaf.  logf is the correctly rounded one,
 * (float)log((double)x); the reference's CUDA logf is documented within 1 ulp
 * of it, which only running the reference could pin further. */
float orc_nmi_ref32(const uint32_t *hist, int H, int64_t N);

/* The same with every logf result moved by -1, 0 or +1 ulp, a fixed choice
 * per argument drawn from `seed` (seed 0: none): what another logf within
 * CUDA's documented 1 ulp could give.  Measures how far the unpinned logf
 * can move a map entry. */
float orc_nmi_ref32_ulp(const uint32_t *hist, int H, int64_t N, uint64_t seed);

/* Checks the correctly rounded logf used above over the floats whose bit
 * patterns lie in [lo_bits, hi_bits] (only the integer-valued ones when
 * integers_only != 0): counts the x whose exact logarithm (logl) lies within
 * rel * |log x| of a float rounding boundary (there a double logarithm with a
 * few ulps of error could round either way), or whose (float)log((double)x)
 * is not the float nearest logl(x).  With rel = 2^-48 the integer counts
 * 1 .. 16,200,281 give 0: each one's logf has one answer that any double log
 * within a few ulps rounds to.  *first_bad (may be NULL) = smallest failing
 * pattern. */
int64_t orc_logf_cr_margin(uint32_t lo_bits, uint32_t hi_bits, double rel, int integers_only,
                           uint32_t *first_bad);

/* The same quantity from exact integer counts, accumulated in double:
 * H(X) = ln N - (1/N) sum c ln c;  NMI = 2(Hx+Hy-Hxy)/(Hx+Hy), NaN -> 0. */
double orc_nmi_f64(const uint32_t *hist, int H, int64_t N);

/* Whole similarity map over n fields (contiguous, stride F): the task list of
 * calculate_histogram_h (JointHistogram.cu:305-330).  multi != 0: fields
 * 0..n-1 against n..2n-1, map[i*n + j] (TSM/VSM).  Writes both the
 * reference-order float map and the double map (either may be NULL).
 * compat_chunks != 0 applies Q16. */
void orc_ism(const float *fields, int n, int64_t F, int H, int multi,
             int compat_chunks, float *map_ref32, double *map_f64, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
