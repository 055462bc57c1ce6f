// Distance-field stage shared by the single-surface entry points and the
// batched ISM pipeline.
#pragma once

#include "iso_internal.hpp"

namespace iso {

// Sample grid of a distance field: (X/ds, Y/ds, Z/ds) samples, z fastest.
struct SampleGrid {
  int X, Y, Z, ds;
  int nx, ny, nz;
};

// Distance fields of the point sets in a pool of ntot points (xyz triples,
// float voxel coordinates or int cells, any order).  Point i belongs to set
// d_seg[i] in [0, nseg) (all in set 0 when d_seg is null); h_counts[s] is the
// size of set s.  The field of the f-th non-empty set (in set order) goes to
// d_df + f*F.  Enqueued on ctx->stream.  normalised: the points are already
// divided by max(X,Y,Z).  coherent: the pool comes from an extraction that
// emits each set's points in runs that follow the volume, so the LBVH may
// take its leaves from those runs instead of sorting every point.
int df_pool(iso_ctx *ctx, const void *d_xyz, bool is_int, int64_t ntot, const uint32_t *d_seg,
            int nseg, const uint32_t *h_counts, const SampleGrid &g, int method, float *d_df,
            bool normalised = false, bool coherent = false);

}  // namespace iso
