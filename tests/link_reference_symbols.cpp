// Link check for the drop-in: one translation unit that references every
// symbol the reference's host library (isosurfaces_lib) takes from its CUDA
// library, declared the way the reference declares them -- the five
// extern "C" functions isosurfaces_cuda/funcs.cpp calls (funcs.h:14-15, 23-24,
// 32-36, 43-44, 53-54) and cudaMarchingCube, which types/src/Isosurface.cpp:263
// calls (hdr/MarchingCube.h:12).  Linked against libisosurfaces_compat.so
// only; never run.
#include <hip/hip_runtime.h>

#include <vector>

typedef struct {
  int x, y, z;
} vec3i;  // utils/typeOperation.h: a typedef'd struct, linkage name vec3i
typedef struct vec4f {
  float x, y, z, w;
} vec4f;

using std::vector;

extern "C" void calculate_distances_h(vector<float3> points, int dimX, int dimY, int dimZ,
                                      int dfDownscale, float *h_distances);
extern "C" long approximate_isosurfaces_h(float *h_data, float isovalue, const int dimX,
                                          const int dimY, const int dimZ, const int scale,
                                          int3 **h_points);
extern "C" void calculate_histogram_h(float **distanceFields, float *minValues, float *maxValues,
                                      const int fieldSize, const unsigned long numFields,
                                      const unsigned long histogramSize, unsigned int *h_jointHist,
                                      unsigned int *h_colSums, unsigned int *h_rowSums,
                                      float *h_simMap, const int simMapSize, const bool multi);
extern "C" void generateHierarchy_h(vector<float3> points, const int dimX, const int dimY,
                                    const int dimZ, float *h_distances, const int numSamples,
                                    const int dfDownscale, const bool sampleGiven,
                                    const unsigned long sampleSize, vector<float3> samplePoints);
extern "C" long findExtrema_h(float *h_data, int3 dimensions, float **h_points);
void cudaMarchingCube(const float &iso_val, const vec3i &dim, float *data, const int &maxVerts,
                      std::vector<vec4f> &ret_points);

int main(int argc, char **) {
  if (argc < 1000) return 0;  // the calls below are only linked
  vector<float3> p;
  float f = 0, *pf = nullptr, **ppf = nullptr;
  int3 *pi = nullptr;
  unsigned u = 0;
  calculate_distances_h(p, 1, 1, 1, 1, &f);
  approximate_isosurfaces_h(&f, 0.f, 2, 2, 2, 1, &pi);
  calculate_histogram_h(ppf, &f, &f, 1, 1, 1, &u, &u, &u, &f, 1, false);
  generateHierarchy_h(p, 1, 1, 1, &f, 1, 1, false, 0, p);
  findExtrema_h(&f, make_int3(1, 1, 1), &pf);
  std::vector<vec4f> out;
  vec3i dim{2, 2, 2};
  int mv = 8;
  cudaMarchingCube(0.f, dim, &f, mv, out);
  return 0;
}
