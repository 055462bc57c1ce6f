// Test driver for the C++ entry points of the reference-interface mirror
// (std::vector arguments cannot be reached from ctypes).  Raw little-endian
// float/int files in and out:
//   compat_driver mc    vol.f32  X Y Z iso        out.f32   (vec4f per vertex)
//   compat_driver df    pts.f32  n X Y Z ds       out.f32   (generateHierarchy_W)
//   compat_driver dfh   pts.f32  n X Y Z ds       out.f32   (generateHierarchy_h, normalised)
//   compat_driver dfi3  pts.i32  n X Y Z ds       out.f32   (generateHierarchyi3_W)
//   compat_driver dist  pts.f32  n nx ny nz ds    out.f32   (calculate_distances_h)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "isosurfaces_compat.hpp"

template <typename T>
static std::vector<T> read_all(const char *path, size_t n) {
  std::vector<T> v(n);
  FILE *f = fopen(path, "rb");
  if (!f || fread(v.data(), sizeof(T), n, f) != n) {
    fprintf(stderr, "cannot read %s\n", path);
    exit(2);
  }
  fclose(f);
  return v;
}

template <typename T>
static void write_all(const char *path, const T *p, size_t n) {
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(p, sizeof(T), n, f) != n) {
    fprintf(stderr, "cannot write %s\n", path);
    exit(2);
  }
  fclose(f);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  if (!strcmp(argv[1], "mc") && argc == 8) {
    vec3i dim{atoi(argv[3]), atoi(argv[4]), atoi(argv[5])};
    float iso = (float)atof(argv[6]);
    auto vol = read_all<float>(argv[2], (size_t)dim.x * dim.y * dim.z);
    std::vector<vec4f> out;
    int maxVerts = dim.x * dim.y * dim.z;
    cudaMarchingCube(iso, dim, vol.data(), maxVerts, out);
    write_all(argv[7], (const float *)out.data(), 4 * out.size());
    printf("%zu\n", out.size());
    return 0;
  }
  if ((!strcmp(argv[1], "df") || !strcmp(argv[1], "dfh") || !strcmp(argv[1], "dfi3")) && argc == 9) {
    long n = atol(argv[3]);
    int X = atoi(argv[4]), Y = atoi(argv[5]), Z = atoi(argv[6]), ds = atoi(argv[7]);
    std::vector<float> df((size_t)(X / ds) * (Y / ds) * (Z / ds));
    if (!strcmp(argv[1], "dfi3")) {
      auto p = read_all<int3>(argv[2], (size_t)n);
      generateHierarchyi3_W(p.data(), X, Y, Z, df.data(), n, 1500, ds);
    } else {
      auto p = read_all<float3>(argv[2], (size_t)n);
      if (!strcmp(argv[1], "df"))
        generateHierarchy_W(p, X, Y, Z, df.data(), 1500, ds);
      else
        generateHierarchy_h(p, X, Y, Z, df.data(), 1500, ds, false, 0, p);
    }
    write_all(argv[8], df.data(), df.size());
    printf("%zu\n", df.size());
    return 0;
  }
  if (!strcmp(argv[1], "dist") && argc == 9) {
    long n = atol(argv[3]);
    int nx = atoi(argv[4]), ny = atoi(argv[5]), nz = atoi(argv[6]), ds = atoi(argv[7]);
    auto p = read_all<float3>(argv[2], (size_t)n);
    std::vector<float> df((size_t)nx * ny * nz);
    calculate_distances_h(p, nx, ny, nz, ds, df.data());
    write_all(argv[8], df.data(), df.size());
    printf("%zu\n", df.size());
    return 0;
  }
  fprintf(stderr, "usage: see header\n");
  return 2;
}
