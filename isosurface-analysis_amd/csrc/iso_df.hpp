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

// Distance fields of S point sets stored end to end in d_xyz (xyz triples,
// float voxel coordinates or int cells); set s spans points
// [d_off[s], d_off[s+1]) (device array of S+1 offsets, every set non-empty)
// and its field goes to d_df + s*F.  Enqueued on ctx->stream, no host sync.
// normalised: the points are already divided by max(X,Y,Z).
int df_segments(iso_ctx *ctx, const void *d_xyz, bool is_int, int64_t ntot, const int64_t *d_off,
                int S, const SampleGrid &g, int method, float *d_df, bool normalised = false);

}  // namespace iso
