// Host-side mirror of the reference's device-library interface
// (isosurfaces_lib/isosurfaces_cuda/funcs.h and hdr/MarchingCube.h), implemented
// on libiso_amd.so.  Same names, argument meanings and ownership rules, so the
// reference's runtime modules (Single, Scripting, MultiVar, ...) can be rebuilt
// against it unchanged:
//   - approximate_isosurfaces_h mallocs *h_points; the caller frees;
//   - h_distances / h_simMap / sums are caller-owned;
//   - jointHist is never written (JointHistogram.cu:397);
//   - errors print a message and exit(EXIT_FAILURE), as checkCudaErrors does.
// One process-wide context on device 0, as the reference uses the default
// device; calls are serialised like the reference's (not re-entrant).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "iso_amd.h"

// Vector types of the reference's utils/typeOperation.h used by this interface.
#ifndef ISO_COMPAT_NO_VEC_TYPES
struct vec3i {
  int x, y, z;
};
struct vec4f {
  float x, y, z, w;
};
#endif

// funcs.h:23-24 -- voxel approximation; returns the number of cells.
extern "C" long approximate_isosurfaces_h(float *h_data, float isovalue, const int dimX,
                                          const int dimY, const int dimZ, const int scale,
                                          int3 **h_points);
// funcs.h:26
long approximate_isosurfaces_W(float *data, float isovalue, const int dimX, const int dimY,
                               const int dimZ, const int scale, int3 **h_points);

// funcs.h:32-36 -- similarity map of numFields (multi: 2*numFields) host fields.
extern "C" void calculate_histogram_h(float **distanceFields, float *minValues, float *maxValues,
                                      const int fieldSize, const unsigned long numFields,
                                      const unsigned long histogramSize, unsigned int *h_jointHist,
                                      unsigned int *h_colSums, unsigned int *h_rowSums,
                                      float *h_simMap, const int simMapSize, const bool multi);

// funcs.h:43-44 -- distance field of points already divided by max(X,Y,Z).
// numSamples (the reference's 1500-leaf pre-sample) only seeds the search
// bound there and cannot change the exact minimum; it is accepted and unused.
// sampleGiven (surface-to-surface queries) is not supported.
extern "C" void generateHierarchy_h(std::vector<float3> points, const int dimX, const int dimY,
                                    const int dimZ, float *h_distances, const int numSamples,
                                    const int dfDownscale, const bool sampleGiven,
                                    const unsigned long sampleSize,
                                    std::vector<float3> samplePoints);
// funcs.h:46-47 -- normalise (p / max(X,Y,Z)) then generateHierarchy_h.
void generateHierarchy_W(std::vector<float3> points, const int dimX, const int dimY, const int dimZ,
                         float *distances, int numSamples, const int dfDownscale);
void generateHierarchyi3_W(int3 *points, const int dimX, const int dimY, const int dimZ,
                           float *distances, const long numPoints, const int numSamples,
                           const int dfDownscale);

// funcs.h:14-15 -- brute-force distance field in voxel units over a
// dimX x dimY x dimZ grid of samples at (x*ds + ds/2, ...) (DistanceField.cu:
// 15-98); h_distances holds dimX*dimY*dimZ floats, caller-owned.  Called by
// the reference's calculateDistances_W (funcs.cpp:71-80).
extern "C" void calculate_distances_h(std::vector<float3> points, int dimX, int dimY, int dimZ,
                                      int dfDownscale, float *h_distances);

// funcs.h:53-54 -- values of the local extrema of a volume in voxel order
// (IsosurfaceExtrema.cu:22-138); mallocs *h_points, the caller frees; returns
// their number.  Called by the reference's findExtrema_W (funcs.cpp:195-198).
extern "C" long findExtrema_h(float *h_data, int3 dimensions, float **h_points);

// hdr/MarchingCube.h:12 -- unique vertices, sorted, w = 1.
void cudaMarchingCube(const float &iso_val, const vec3i &dim, float *data, const int &maxVerts,
                      std::vector<vec4f> &ret_points);

// The context this mirror uses (created on first use).
iso_ctx *isosurfaces_compat_context();
