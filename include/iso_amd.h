/*
 * iso_amd.h -- C ABI of libiso_amd.so, the MI355X (gfx950) implementation of
 * the isosurface-similarity-map (ISM) hot path of mimre25/isosurface-analysis:
 * surface extraction at k isovalues -> one distance field per isosurface ->
 * pairwise joint histograms and normalised mutual information.
 *
 * Plain C: POD arguments, int status codes, no exit(), no C++ types.
 * Reference citations are relative to isosurfaces_lib/ of the reference.
 *
 * Conventions
 *   - Every entry point takes an iso_ctx (one HIP device + one stream +
 *     a reusable device workspace).  Calls on one context are serialised on
 *     its stream; a context must not be used from two host threads at once.
 *   - d_ prefixed arguments are device pointers (HBM of the context's device),
 *     h_ prefixed ones are host pointers.  Entry points that return counts to
 *     the host synchronise the context's stream before returning.
 *   - Volumes are float32.  The reference's image buffer `img` is z-fastest:
 *     img[(x*Y + y)*Z + z] (types/src/Isosurface.cpp:185-200).
 *   - Distance-field sample s of grid (X/ds, Y/ds, Z/ds) is z-fastest
 *     (isosurfaces_cuda/src/BVH.cu:393-406).
 */
#ifndef ISO_AMD_H
#define ISO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----------------------------------------------------- */
#define ISO_OK 0
#define ISO_ERR_INVALID_ARG 1
#define ISO_ERR_HIP 2
#define ISO_ERR_OUT_OF_MEMORY 3
#define ISO_ERR_CAPACITY 4 /* output buffer too small; the count out-arg holds the size needed */
#define ISO_ERR_UNSUPPORTED 5
#define ISO_ERR_EMPTY 6    /* distance field of an empty point set */

/* ---- reference-compatibility switches (bitmask `flags`) ---------------- */
/* Marching cubes reads the buffer x-fastest with dims (gx,gy,gz), samples
 * clamped corners so border cells emit vertices at coordinate = dim (Q6),
 * and classifies only the 128*grid.x*grid.y cells of the reference launch
 * (Q4).  isosurfaces_cuda/src/MarchingCube.cu:59-124, MarchingCube.cpp:599-613.
 * Without it ("fixed"): the buffer is the z-fastest image with dims (X,Y,Z),
 * all (X-1)(Y-1)(Z-1) interior cells, vertices in (x,y,z) image coordinates. */
#define ISO_COMPAT_MC (1u << 0)
/* Joint histograms skip the final 1,000,000 voxels when F is a positive
 * multiple of 1,000,000, as the reference's chunk loop does (Q16).
 * isosurfaces_cuda/src/JointHistogram.cu:372-393. */
#define ISO_COMPAT_HIST_CHUNKS (1u << 1)
/* NMI in the reference's float arithmetic and order: per map entry (i, j)
 * the fused multiply-add chains hXY -= c*logf(c) over the nonzero bins of the
 * (i, j) histogram (rows outer), hX and hY over the column and row sums, then
 * h/N + logf(N) and 2(hX+hY-hXY)/(hX+hY) in float, NaN -> 0
 * (calculateMutualInformation, isosurfaces_cuda/src/JointHistogram.cu:24-60).
 * logf is the correctly rounded one (the reference's CUDA logf is within 1
 * ulp of it).  Entries (i, j) and (j, i) can differ in the last bits, as in
 * the reference.  Without it: the same quantity from the exact counts in
 * double, rounded once to float (symmetric map). */
#define ISO_COMPAT_NMI32 (1u << 2)
#define ISO_COMPAT_ALL (ISO_COMPAT_MC | ISO_COMPAT_HIST_CHUNKS | ISO_COMPAT_NMI32)
/* Marching cubes: return the unique vertices sorted lexicographically (x,y,z),
 * as the reference's removeDuplicates does (thrustWrapper.cu:44-56).
 * Without it the order is unspecified (the set is identical). */
#define ISO_MC_SORTED (1u << 8)
/* iso_ism_batch: run the pair stage (queue reset + the persistent histogram
 * + NMI launch) as a captured HIP graph.  The graph is captured on the first
 * call with a given pair-task layout and replayed by later calls with the
 * same layout (same volumes sizes, isosurface counts and bins); the maps are
 * identical to the eager launch. */
#define ISO_BATCH_GRAPH (1u << 9)

/* ---- distance-field methods ------------------------------------------- */
#define ISO_DF_LBVH 0  /* Morton sort + Karras tree + nearest-point traversal */
#define ISO_DF_BRUTE 1 /* every (sample, point) pair; the LBVH equals it bit for bit */

/* ---- context ----------------------------------------------------------- */
typedef struct iso_ctx iso_ctx;

/* Creates a context on HIP device `device` with its own stream.
 * (The reference keeps process globals on the default stream of device 0:
 * isosurfaces_cuda/src/MarchingCube.cpp:532-569.) */
int iso_ctx_create(iso_ctx **out, int device);
int iso_ctx_destroy(iso_ctx *ctx);
int iso_ctx_synchronize(iso_ctx *ctx);
/* The context's hipStream_t, as void*. */
void *iso_ctx_stream(iso_ctx *ctx);
/* Text of the last error raised on this context ("" if none). */
const char *iso_ctx_last_error(const iso_ctx *ctx);
/* The autotuner's choices on this context so far, one line per shape class:
 * the launch variants' times on the first call of that class and the variant
 * kept (every variant computes the same result).  ISO_AUTOTUNE=0 in the
 * environment keeps default variants.  No counterpart in the reference. */
const char *iso_ctx_tuning(const iso_ctx *ctx);
const char *iso_status_string(int status);

/* Device-memory helpers so a host program needs nothing but this library. */
int iso_device_alloc(iso_ctx *ctx, void **d_ptr, size_t bytes);
int iso_device_free(iso_ctx *ctx, void *d_ptr);
int iso_copy_to_device(iso_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int iso_copy_to_host(iso_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int iso_device_memset(iso_ctx *ctx, void *d_dst, int value, size_t bytes);
/* Device-to-device copy; 16-byte aligned sizes go through a plain copy
 * kernel that moves exactly `bytes` each way (used to calibrate counters). */
int iso_device_copy(iso_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);

/* ---- volume ingest and isovalues ---------------------------------------- */
/* Raw x-fastest volume (the `.dat` file as read) -> z-fastest image
 * d_img[(x*Y+y)*Z+z] = d_file[x + y*X + z*X*Y], and the extremes of the
 * values inside [cfg_min, cfg_max] (from 100000 / -100000).  Replaces
 * BinaryFloatReader::read + Isosurface::loadFile + createImgPtr
 * (filehandler/src/BinaryFloatReader.cpp:29-49, types/src/Isosurface.cpp:
 * 32-76, 185-200).  d_file and d_img must not overlap. */
int iso_volume_ingest(iso_ctx *ctx, const float *d_file, int X, int Y, int Z, float cfg_min,
                      float cfg_max, float *d_img, float *h_real_min, float *h_real_max);

/* k isovalues by 65536-bin histogram equalisation over [real_min, real_max]
 * (Single::compute_isovalues with uniform = true, runtime/src/Single.cpp:
 * 286-392): device histogram, host walk over the bins. */
int iso_equalized_isovalues(iso_ctx *ctx, const float *d_img, int64_t n, float real_min,
                            float real_max, int k, float *h_isos);

/* ---- surface extraction ------------------------------------------------ */
typedef struct {
  int64_t active_cells; /* classified cells with at least one vertex */
  int64_t total_verts;  /* sum of the case's vertex count over classified cells */
  int64_t triangles;    /* total_verts / 3 */
  int64_t unique_verts; /* vertices after exact-duplicate removal */
} iso_mc_stats;

/* Marching cubes at one isovalue; writes the unique vertices (xyz triples,
 * voxel coordinates) to d_xyz.  Replaces cudaMarchingCube
 * (isosurfaces_cuda/hdr/MarchingCube.h:12, src/MarchingCube.cpp:588-710).
 * With ISO_COMPAT_MC, (gx,gy,gz) are the dims the reference passes
 * (vec3i(X,Y,Z), types/src/Isosurface.cpp:261-263) and d_vol is `img`.
 * Returns ISO_ERR_CAPACITY (and *h_count = unique count) if cap is too small.
 * h_stats may be NULL. */
int iso_mc_extract(iso_ctx *ctx, const float *d_vol, int gx, int gy, int gz, float iso,
                   uint32_t flags, float *d_xyz, int64_t cap, int64_t *h_count,
                   iso_mc_stats *h_stats);

/* Voxel approximation at one isovalue: integer corners (x,y,z) of the cells
 * of the z-fastest image that have a corner > iso and a corner < iso, in
 * x-major / z-fastest order.  Replaces approximate_isosurfaces_h
 * (isosurfaces_cuda/funcs.h:23-24, src/IsosurfaceApproximation.cu:159-212). */
int iso_approx_extract(iso_ctx *ctx, const float *d_img, int X, int Y, int Z, float iso,
                       int32_t *d_xyz, int64_t cap, int64_t *h_count);

/* ---- distance fields ---------------------------------------------------- */
/* Distance field of n points over the (X/ds, Y/ds, Z/ds) sample grid:
 * d_df[s] = sqrtf(min_p d2(s,p)) (d2 itself when n == 1), points normalised
 * by max(X,Y,Z) and samples per axis as the reference does.  Replaces
 * generateHierarchy_h / generateHierarchy_W / generateHierarchyi3_W
 * (isosurfaces_cuda/funcs.h:43-47, funcs.cpp:209-233, src/BVH.cu:814-875).
 * _f3: float xyz in voxel coordinates (marching cubes); _i3: int xyz cells
 * (approximation).  method = ISO_DF_LBVH or ISO_DF_BRUTE. */
int iso_df_points_f3(iso_ctx *ctx, const float *d_xyz, int64_t n, int X, int Y, int Z,
                     int ds, int method, float *d_df);
int iso_df_points_i3(iso_ctx *ctx, const int32_t *d_xyz, int64_t n, int X, int Y, int Z,
                     int ds, int method, float *d_df);
/* Same, for points already divided by max(X,Y,Z) -- what generateHierarchy_h
 * receives from its _W wrappers (isosurfaces_cuda/funcs.cpp:209-233). */
int iso_df_points_normalized(iso_ctx *ctx, const float *d_xyz, int64_t n, int X, int Y, int Z,
                             int ds, int method, float *d_df);

/* Brute-force distance field in voxel units: d_df[s] = sqrtf(min_p d2(s,p))
 * over a grid of nx*ny*nz samples (z fastest) at (x*ds + ds/2, y*ds + ds/2,
 * z*ds + ds/2), points taken as given (float xyz, not normalised).  The
 * minimum starts at 1e38 (no points: sqrtf(1e38)) and the points are taken in
 * runs of 2,000,000, each run starting from the square of the previous run's
 * root, as the reference launches them.  Replaces calculate_distances_h
 * (isosurfaces_cuda/funcs.h:14-15, src/DistanceField.cu:15-98), which the
 * reference's funcs.cpp:71-80 (calculateDistances_W) calls. */
int iso_df_voxel_units(iso_ctx *ctx, const float *d_xyz, int64_t n, int nx, int ny, int nz,
                       int ds, float *d_df);

/* ---- volume extrema ----------------------------------------------------- */
/* Values of the voxels of d_data (X*Y*Z floats) that are local extrema, in
 * voxel order (z fastest): a voxel's value and its neighbours' are read at
 * x + y*Y + z*Y*Z (the reference's own index); neighbours are x-1..x+3,
 * y-1..y+3, z-1..z+3 inside the grid, the voxel included; a voxel is dropped
 * when more than 5 of them equal it or when some are lower and some higher,
 * and values at or below -1e10 are dropped too.  ISO_ERR_CAPACITY (and
 * *h_count = the number of values) if cap is too small; ISO_ERR_UNSUPPORTED
 * for dims whose reference index leaves the volume (the reference reads past
 * its buffer there).  Replaces findExtrema_h (isosurfaces_cuda/funcs.h:53-54,
 * src/IsosurfaceExtrema.cu:22-138), which funcs.cpp:195-198 (findExtrema_W)
 * calls. */
int iso_find_extrema(iso_ctx *ctx, const float *d_data, int X, int Y, int Z, float *d_out,
                     int64_t cap, int64_t *h_count);

/* ---- joint histograms + mutual information ------------------------------ */
/* Per-field (min, max) over F values; max starts at FLT_MIN as in
 * DistanceField::getInterval (types/src/DistanceField.cpp:159-200). */
int iso_field_interval(iso_ctx *ctx, const float *d_fields, int nfields, int64_t F,
                       float *d_mins, float *d_maxs);

/* Similarity map over n fields stored contiguously (field i at d_fields+i*F).
 * Plain: d_map[i*n+j] = NMI of the H-bin joint histogram of fields (i, j),
 * all n*n entries.  multi != 0: d_fields holds 2n fields and
 * d_map[i*n+j] pairs field i with field n+j (TSM/VSM).  Bins follow
 * findBucket_d with the pair range [min(min_i,min_j), max(max_i,max_j)];
 * counts are exact; NMI follows the reference's float order with
 * ISO_COMPAT_NMI32 in flags, else it is evaluated from the exact counts in
 * double and rounded to float.  Optional outputs (NULL to skip), reference layout:
 * d_joint[(i*n+j)*H*H + row*H + col], d_rows/d_cols[(i*n+j)*H + b].
 * Replaces calculate_histogram_h (isosurfaces_cuda/funcs.h:32-36,
 * src/JointHistogram.cu:266-420).  H in [1, 256]. */
int iso_hist_nmi(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                 uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                 uint32_t *d_joint, uint32_t *d_rows, uint32_t *d_cols);

/* iso_hist_nmi without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, same results as iso_hist_nmi (the same kernel; the
 * default launch shape, no autotuning).  Its working memory is a
 * 64 MiB array per device, held until the process ends; pairs go through
 * in launches of as many as it holds (127 pairs at 256 bins).  Rejects
 * H > 256 and more than 65535 fields (ISO_ERR_UNSUPPORTED).  Calls that
 * overlap on one device share that array and race: serialise them. */
int iso_hist_nmi_noalloc(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                         uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                         uint32_t *d_joint, uint32_t *d_rows, uint32_t *d_cols);

/* ---- batched ISM pipeline ---------------------------------------------- */
#define ISO_EXTRACT_MC 0     /* marching cubes (the reference's `script` module) */
#define ISO_EXTRACT_APPROX 1 /* voxel approximation (the reference's `single` module) */

typedef struct {
  int n_nonempty;       /* isosurfaces with at least one point (= map size m) */
  int64_t total_points; /* points over all isosurfaces */
  double ms_extract;    /* stage times on the device (HIP events) */
  double ms_df;
  double ms_hist;
  int64_t graph_replays; /* iso_ism_batch with ISO_BATCH_GRAPH: earlier replays of
                            the captured pair stage (0 when this call captured it) */
  double ms_walk;        /* iso_ism_volume with the LBVH: the tree-walk kernel alone
                            (part of ms_df) */
} iso_ism_stats;

/* The whole isosurface similarity map of one volume in one call: extraction
 * at k isovalues, a distance field per non-empty isosurface on the
 * (X/ds,Y/ds,Z/ds) grid, then the H-bin NMI of every pair.  Writes the m x m
 * map (row-major, m = non-empty isosurfaces) to d_map (room for k*k floats),
 * the caller indices of the non-empty isovalues to h_ids (room for k), and,
 * if d_fields != NULL, the m fields (m*F floats) to d_fields.  With
 * ISO_EXTRACT_MC, d_vol is read as iso_mc_extract reads it (ISO_COMPAT_MC in
 * flags); with ISO_EXTRACT_APPROX it is the z-fastest image.
 * Replaces the loop of Single::run (runtime/src/Single.cpp:193-256) over
 * Isosurface::calculateSurfacePoints, DistanceField::calculateDistanceField
 * and calculate_histogram_W.  h_stats may be NULL. */
int iso_ism_volume(iso_ctx *ctx, const float *d_vol, int X, int Y, int Z, const float *h_isos,
                   int k, int extractor, int ds, int H, int df_method, uint32_t flags,
                   float *d_map, float *d_fields, int *h_ids, iso_ism_stats *h_stats);

/* One volume of a batch: inputs as for iso_ism_volume; d_map (room for k*k
 * floats) and h_ids (room for k) are outputs, n_nonempty is set on return. */
typedef struct {
  const float *d_vol;
  int X, Y, Z;
  const float *h_isos;
  int k;
  float *d_map;
  int *h_ids;
  int n_nonempty;
} iso_ism_job;

/* The ISMs of several volumes (any sizes) in one call: extraction and fields
 * volume by volume, then every pair of every ISM in one persistent histogram
 * + NMI launch that pulls pair tasks from a device queue, largest fields
 * first.  Each map equals what iso_ism_volume returns for the same job.
 * Replaces running the reference's `single` process once per volume
 * (runtime/src/Single.cpp:193-256, one histogram launch per 1M-voxel chunk per
 * volume).  h_stats (may be NULL): ms_df covers extraction + fields of all
 * volumes, ms_hist the pair launch. */
int iso_ism_batch(iso_ctx *ctx, iso_ism_job *jobs, int njobs, int extractor, int ds, int H,
                  int df_method, uint32_t flags, iso_ism_stats *h_stats);

/* ---- representatives ---------------------------------------------------- */
/* Selects cnt representative isosurfaces from an m x m similarity map with the
 * reference's procedure (SimilarityMap::findRepresentativeIsovalues,
 * types/src/SimilarityMap.cpp:88-227).  h_ids (may be NULL) maps map rows to
 * caller ids.  Writes the chosen ids ascending (the order of the reference's
 * distancefield-<n>.df files, runtime/src/Single.cpp:476-545) and, if not
 * NULL, their priorities when chosen; *h_count = min(cnt, m).  Host only. */
int iso_select_representatives(const float *h_map, int m, const int *h_ids, int cnt,
                               int *h_out_ids, float *h_out_priority, int *h_count);

/* The same selection in the order the representatives are taken (the keys
 * 0..cnt-1 of the reference's result map, SimilarityMap.cpp:162-199): caller
 * ids, map rows (may be NULL) and priorities when taken (may be NULL). */
int iso_select_representatives_ordered(const float *h_map, int m, const int *h_ids, int cnt,
                                       int *h_sel_ids, int *h_sel_map_ids,
                                       float *h_sel_priority, int *h_count);

/* ---- serving --------------------------------------------------------------- */
/* A request loop over HBM-resident volumes answering ISM, TSM (one variable,
 * two time steps) and VSM (two variables, one time step) requests; the cross
 * maps pair the representative fields of two volumes as the reference's
 * `multi` module does (runtime/src/MultiVar.cpp:168-249, multi=true task list
 * of JointHistogram.cu:309-318).  No counterpart exists in the reference's
 * process-per-map runtime; this is the serving form of the same hot path. */
#define ISO_REQ_ISM 0
#define ISO_REQ_TSM 1
#define ISO_REQ_VSM 2

typedef struct {
  int device;
  int slots;          /* concurrent requests on the device (one stream each) */
  size_t mem_budget;  /* bytes of cached fields + in-flight fields; the server
                         allocates this many bytes of field blocks (at least two,
                         at most one per volume) when it starts, so no request
                         allocates or frees device memory */
  int k;              /* isovalues per ISM (evenly spaced inside the volume range) */
  int ds;             /* distance-field downscale */
  int H_ism;          /* histogram bins of an ISM */
  int H_multi;        /* histogram bins of a cross map (reference: 128) */
  int reps;           /* representatives per volume (reference: 16) */
  int extractor;      /* ISO_EXTRACT_MC or ISO_EXTRACT_APPROX */
  int df_method;
  uint32_t flags;
} iso_server_config;

typedef struct {
  int kind; /* ISO_REQ_* */
  int var1, step1;
  int var2, step2; /* second volume of TSM / VSM */
} iso_request;

typedef struct {
  int status;
  int m;    /* map is m x m */
  int slot; /* worker slot that ran it */
  double t_submit, t_start, t_done; /* steady-clock seconds */
  int computed; /* ISMs this request computed (0: every volume it used was cached) */
  int64_t computed_pairs; /* isovalue pairs of those ISMs' maps, m(m-1)/2 each */
} iso_request_info;

typedef struct {
  int64_t submitted, completed, ism_computed, cache_hits, evictions;
  size_t cache_bytes;
  int resident_volumes;
  /* launch-shape measurements the autotuner took inside requests (0: it
   * measures only during iso_server_start), and launches whose shape class
   * was not measured there and took the nearest measured choice */
  int64_t tunings_in_requests, untuned_shapes;
} iso_server_stats;

typedef struct iso_server iso_server;
int iso_server_create(iso_server **out, const iso_server_config *cfg);
/* Uploads a z-fastest volume (all volumes share one size); before start. */
int iso_server_add_volume(iso_server *s, int var, int step, const float *h_img, int X, int Y,
                          int Z);
int iso_server_start(iso_server *s);
/* Non-blocking; the arrival time is stamped here. */
int iso_server_submit(iso_server *s, const iso_request *req, uint64_t *ticket);
/* Blocks until the request is done.  ISM: the m x m map and the m non-empty
 * isovalue indices; TSM/VSM: the n x n cross map and the 2n representative
 * isovalue indices (first volume, then second). */
int iso_server_wait(iso_server *s, uint64_t ticket, float *h_map, int map_cap, int *h_ids,
                    int ids_cap, iso_request_info *info);
int iso_server_get_stats(iso_server *s, iso_server_stats *out);
int iso_server_destroy(iso_server *s);

/* ---- workload generators (host, deterministic, z-fastest image) ---------- */
/* Multivariate time-varying set: Gaussian blobs advected through a periodic
 * box; var 0 = f, var 1 = exp(f), var 2 = |grad f|^2. */
int iso_gen_timeseries(float *h_img, int X, int Y, int Z, int var, int step, int nblobs,
                       uint64_t seed);
/* Sum of nblobs isotropic Gaussians (centre uniform in the box, sigma in
 * [4,12], amplitude in [0.5,1.5]) plus uniform noise in [-noise, +noise]. */
int iso_gen_blobs(float *h_img, int X, int Y, int Z, int nblobs, float noise, uint64_t seed);
/* clamp(0.5 + 0.5 tanh((x - x0(y,z))/w) + 0.1 * valuenoise3(x,y,z), 0, 1),
 * x0 = X/2 + A sin(2 pi y/ly) sin(2 pi z/lz), w = 8, A = 20, ly = Y/2, lz = Z/2,
 * value noise of 3 octaves (lattice 32, 16, 8 voxels). */
int iso_gen_front(float *h_img, int X, int Y, int Z, uint64_t seed);
/* sum of 4 products sin(a x) sin(b y) sin(c z), frequencies in [0.05, 0.4]. */
int iso_gen_sines(float *h_img, int X, int Y, int Z, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
