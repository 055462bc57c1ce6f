// isosurfaces_amd: the analysis runtime's modules on the HIP library.
//
//   isosurfaces_amd single <SingleConfig.json>
//       one volume, voxel approximation, 256 equalised isovalues
//       (runtime/main.cpp:192-252 building Settings(HIST_SIZE, MAP_SIZE,
//       MAP_SIZE, ..., approximation=true) and running Single::run)
//   isosurfaces_amd script <ScriptingInput.json>
//       every variable x time step of every entry through the same flow with
//       marching cubes (runtime/src/Scripting.cpp:22-236, SINGLE build)
//   isosurfaces_amd multi <MultiVarConfig.json>
//       TSM/VSM cross maps of the saved representative distance fields
//       (runtime/src/MultiVar.cpp:76-249)
//   isosurfaces_amd distance <SingleConfig.json>
//       isovalue equalisation of 17 isosurfaces by their NMI distances
//       (runtime/main.cpp:281-343, runtime/src/DistanceEqualization.cpp)
//
// Outputs follow the reference's layout: <out>/<var>-<t>/similarityMap-<app>/
// map.dat (m x m float32), <out>/<var>-<t>/distancefields/distancefield-<n>.df
// (the n-th representative in ascending isovalue index, F float32), an
// isosurfaces-<app>/ folder, runtimes.csv, and the run-information file
// (meta.json) with the field order of the reference's types.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "iso_amd.h"
#include "json_lite.hpp"

namespace fs = std::filesystem;

namespace {

// constants of the reference's runtime (runtime/hdr/globals.h:12-32)
constexpr int kHistSize = 128;  // HIST_SIZE
constexpr int kMapSize = 256;   // MAP_SIZE: isovalues per volume
constexpr int kScale = 1;       // SCALE
constexpr int kReps = 16;       // MULTI_MAP_DF_NUMBER

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---- device context, created on first use (skipped runs need no GPU) -----
struct Gpu {
  iso_ctx *ctx = nullptr;
  ~Gpu() {
    if (ctx) iso_ctx_destroy(ctx);
  }
  iso_ctx *get() {
    if (!ctx) check(iso_ctx_create(&ctx, 0), "iso_ctx_create");
    return ctx;
  }
  void check(int st, const char *what) {
    if (st != ISO_OK) {
      std::string msg = std::string(what) + ": " + iso_status_string(st);
      if (ctx) msg += std::string(" (") + iso_ctx_last_error(ctx) + ")";
      throw Failure(msg);
    }
  }
};

struct DeviceBuf {
  Gpu &g;
  void *p = nullptr;
  DeviceBuf(Gpu &gpu, size_t bytes) : g(gpu) { g.check(iso_device_alloc(g.get(), &p, bytes), "alloc"); }
  ~DeviceBuf() {
    if (p) iso_device_free(g.ctx, p);
  }
  template <class T> T *as() const { return (T *)p; }
};

// ---- files -----------------------------------------------------------------
std::string read_text(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Failure("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

void write_bytes(const std::string &path, const void *data, size_t bytes, bool append = false) {
  std::ofstream f(path, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
  if (!f) throw Failure("cannot write " + path);
  f.write((const char *)data, (std::streamsize)bytes);
  if (!f) throw Failure("write failed: " + path);
}

std::vector<float> read_floats(const std::string &path, size_t n) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw Failure("cannot open " + path);
  const size_t have = (size_t)f.tellg();
  if (have < n * sizeof(float))
    throw Failure(path + ": " + std::to_string(have) + " bytes, need " +
                  std::to_string(n * sizeof(float)));
  std::vector<float> v(n);
  f.seekg(0);
  f.read((char *)v.data(), (std::streamsize)(n * sizeof(float)));
  return v;
}

std::string fmt(const char *f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char *f, ...) {
  va_list ap;
  va_start(ap, f);
  char buf[4096];
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

// ---- run information (DAO/hdr/RunInformation.h and friends) -------------
jl::Value new_run_info(int X, int Y, int Z, int ds) {
  jl::Value r = jl::Value::object();
  jl::Value dims = jl::Value::array();
  for (int d : {X, Y, Z}) dims.arr.push_back(jl::Value::of_int(d));
  r.set("dimensions", dims);
  r.set("variables", jl::Value::array());
  r.set("dfDownscale", jl::Value::of_int(ds));
  r.set("volumes", jl::Value::array());
  r.set("similarityMaps", jl::Value::array());
  return r;
}

// index of a variable name, appended when new (RunInformation.cpp addVolumeInformation)
int var_index(jl::Value &run, const std::string &name) {
  jl::Value &vars = run.at("variables");
  for (size_t i = 0; i < vars.arr.size(); ++i)
    if (vars.arr[i].as_str() == name) return (int)i;
  vars.arr.push_back(jl::Value::of(name));
  return (int)vars.arr.size() - 1;
}

void add_similarity_map(jl::Value &run, const std::string &v1, const std::string &v2, int t1,
                        int t2, const std::string &file) {
  jl::Value m = jl::Value::object();
  m.set("var1", jl::Value::of_int(var_index(run, v1)));
  m.set("t1", jl::Value::of_int(t1));
  m.set("var2", jl::Value::of_int(var_index(run, v2)));
  m.set("t2", jl::Value::of_int(t2));
  m.set("filename", jl::Value::of(file));
  run.at("similarityMaps").arr.push_back(m);
}

struct RepInfo {
  int value_id;
  float isovalue;
  int rep_id;
  float importance;
  std::string filename;
  int map_id;
};

// ---- single / script ------------------------------------------------------
struct SingleParams {
  std::string file_name;  // without the .dat extension
  int X = 0, Y = 0, Z = 0;
  float min_value = 0, max_value = 0;
  std::string output_folder, variable_name, json_file;
  int timestep = 0, ds = 1;
  bool approximation = true;
};

// Single::checkIfAlreadyProcessed (Single.cpp:115-140)
bool already_processed(const std::string &json_path, const std::string &var, int ts) {
  if (!fs::exists(json_path)) return false;
  const jl::Value run = jl::parse(read_text(json_path));
  const jl::Value &vars = run.at("variables");
  for (const jl::Value &v : run.at("volumes").arr) {
    const long long vi = v.at("variable").as_int();
    if (vi >= 0 && vi < (long long)vars.arr.size() && vars.arr[(size_t)vi].as_str() == var &&
        v.at("timestep").as_int() == ts)
      return true;
  }
  return false;
}

int run_single(Gpu &gpu, const SingleParams &p) {
  using clock = std::chrono::steady_clock;
  const std::string out = p.output_folder + "/";
  const std::string json_path = out + p.json_file;
  if (already_processed(json_path, p.variable_name, p.timestep)) {
    std::printf("ts %d already processed... skipping\n", p.timestep);
    return 0;
  }
  const auto t0 = clock::now();
  // Single::createFolder (Single.cpp:399-429)
  const std::string approx = p.approximation ? "-approx" : "";
  const std::string scale = p.approximation ? "-scale_" + std::to_string(kScale) : "";
  const std::string appendix = std::to_string(kMapSize) + "-" + approx + scale;
  const std::string var_folder = out + p.variable_name + "-" + std::to_string(p.timestep);
  const std::string iso_folder = var_folder + "/isosurfaces-" + appendix + "/";
  const std::string sim_folder = var_folder + "/similarityMap-" + appendix + "/";
  const std::string df_folder = var_folder + "/distancefields/";
  for (const std::string &d : {iso_folder, sim_folder, df_folder}) fs::create_directories(d);

  const int X = p.X, Y = p.Y, Z = p.Z, ds = p.ds;
  if (X < 2 || Y < 2 || Z < 2 || ds < 1 || X < ds || Y < ds || Z < ds)
    throw Failure("bad dimensions or dfDownscale");
  const size_t N = (size_t)X * Y * Z;
  const int64_t F = (int64_t)(X / ds) * (Y / ds) * (Z / ds);
  const std::string fname = fs::path(p.file_name).filename().string();

  // Isosurface::loadFile (Isosurface.cpp:32-76): x-fastest file -> z-fastest image
  std::vector<float> file = read_floats(p.file_name + ".dat", N);
  iso_ctx *ctx = gpu.get();
  DeviceBuf d_img(gpu, N * 4);
  float real_min = 0, real_max = 0;
  {
    DeviceBuf d_file(gpu, N * 4);
    gpu.check(iso_copy_to_device(ctx, d_file.p, file.data(), N * 4), "copy");
    gpu.check(iso_volume_ingest(ctx, d_file.as<float>(), X, Y, Z, p.min_value, p.max_value,
                                d_img.as<float>(), &real_min, &real_max),
              "iso_volume_ingest");
  }
  file.clear();
  file.shrink_to_fit();
  // Single::compute_isovalues (Single.cpp:286-392), uniform = histogram equalisation
  std::vector<float> isos(kMapSize);
  gpu.check(iso_equalized_isovalues(ctx, d_img.as<float>(), (int64_t)N, real_min, real_max,
                                    kMapSize, isos.data()),
            "iso_equalized_isovalues");
  // the loop of Single::run (Single.cpp:193-256): all isovalues in one call
  DeviceBuf d_map(gpu, sizeof(float) * kMapSize * kMapSize);
  DeviceBuf d_fields(gpu, sizeof(float) * (size_t)F * kMapSize);
  std::vector<int> ids(kMapSize);
  iso_ism_stats st;
  gpu.check(iso_ism_volume(ctx, d_img.as<float>(), X, Y, Z, isos.data(), kMapSize,
                           p.approximation ? ISO_EXTRACT_APPROX : ISO_EXTRACT_MC, ds, kHistSize,
                           ISO_DF_LBVH, ISO_COMPAT_ALL, d_map.as<float>(), d_fields.as<float>(),
                           ids.data(), &st),
            "iso_ism_volume");
  const int m = st.n_nonempty;
  std::printf("%s t=%d: %d of %d isosurfaces non-empty, %lld points\n", p.variable_name.c_str(),
              p.timestep, m, kMapSize, (long long)st.total_points);
  std::vector<float> map((size_t)m * m);
  if (m > 0)
    gpu.check(iso_copy_to_host(ctx, map.data(), d_map.p, map.size() * 4), "copy map");
  // Single::printSimilarityMap -> SimilarityMap::save (raw float32, row-major)
  write_bytes(sim_folder + "map.dat", map.data(), map.size() * 4);

  // SimilarityMap::findRepresentativeIsovalues + Single::printResults
  std::vector<RepInfo> reps;
  if (m > 0) {
    std::vector<int> sel(kReps), rows(kReps);
    std::vector<float> pri(kReps);
    int cnt = 0;
    gpu.check(iso_select_representatives_ordered(map.data(), m, ids.data(), kReps, sel.data(),
                                                 rows.data(), pri.data(), &cnt),
              "iso_select_representatives_ordered");
    std::vector<int> ascending(sel.begin(), sel.begin() + cnt);
    std::sort(ascending.begin(), ascending.end());
    // every pick after the first carries the first pick's isovalue in its
    // heap entry (SimilarityMap.cpp:186): the second number of the file name
    const float carried = isos[(size_t)sel[0]];
    std::vector<float> field((size_t)F);
    for (int r = 0; r < cnt; ++r) {
      const float cur = isos[(size_t)sel[r]];
      reps.push_back(RepInfo{sel[r], cur, r, pri[r],
                             fmt("%s%s-%04d-%0.2f-%0.2f.dat", iso_folder.c_str(), fname.c_str(), r,
                                 (double)cur, (double)(r == 0 ? cur : carried)),
                             rows[r]});
      const int rank = (int)(std::lower_bound(ascending.begin(), ascending.end(), sel[r]) -
                             ascending.begin());
      gpu.check(iso_copy_to_host(ctx, field.data(), d_fields.as<float>() + (size_t)rows[r] * F,
                                 (size_t)F * 4),
                "copy field");
      write_bytes(df_folder + "distancefield-" + std::to_string(rank) + ".df", field.data(),
                  (size_t)F * 4);
    }
  }
  const double secs = std::chrono::duration<double>(clock::now() - t0).count();
  {
    const std::string line = p.variable_name + fname + "," +
                             fmt("%.3f,%.3f,%.3f,%.6f\n", st.ms_extract, st.ms_df, st.ms_hist, secs);
    write_bytes(p.output_folder + "/runtimes.csv", line.data(), line.size(), true);
  }
  // Single::saveRunInfo (Single.cpp:599-633)
  jl::Value run = fs::exists(json_path) ? jl::parse(read_text(json_path))
                                        : new_run_info(X, Y, Z, ds);
  {
    jl::Value vol = jl::Value::object();
    vol.set("variable", jl::Value::of_int(var_index(run, p.variable_name)));
    vol.set("timestep", jl::Value::of_int(p.timestep));
    vol.set("numIsovalues", jl::Value::of_int(m));
    jl::Value ra = jl::Value::array();
    for (const RepInfo &r : reps) {
      jl::Value o = jl::Value::object();
      o.set("valueId", jl::Value::of_int(r.value_id));
      o.set("isovalue", jl::Value::of_float(r.isovalue));
      o.set("repId", jl::Value::of_int(r.rep_id));
      o.set("importance", jl::Value::of_float(r.importance));
      o.set("filename", jl::Value::of(r.filename));
      o.set("mapId", jl::Value::of_int(r.map_id));
      ra.arr.push_back(o);
    }
    vol.set("representatives", ra);
    vol.set("runtime", jl::Value::of_float(secs));
    run.at("volumes").arr.push_back(vol);
  }
  add_similarity_map(run, p.variable_name, p.variable_name, p.timestep, p.timestep,
                     sim_folder + "map.dat");
  const std::string text = jl::dump(run);
  write_bytes(json_path, text.data(), text.size());
  return 0;
}

SingleParams single_from_config(const jl::Value &c) {
  SingleParams p;
  p.file_name = c.at("fileName").as_str();
  const jl::Value &d = c.at("dimensions");
  if (d.arr.size() != 3) throw Failure("dimensions must have 3 entries");
  p.X = (int)d.arr[0].as_int();
  p.Y = (int)d.arr[1].as_int();
  p.Z = (int)d.arr[2].as_int();
  p.min_value = (float)c.at("minValue").as_num();
  p.max_value = (float)c.at("maxValue").as_num();
  p.output_folder = c.at("outputFolder").as_str();
  p.variable_name = c.at("variableName").as_str();
  p.timestep = (int)c.at("timestep").as_int();
  p.json_file = c.at("jsonFile").as_str();
  p.ds = (int)c.at("dfDownscale").as_int();
  p.approximation = true;
  return p;
}

// fileFormatString with three strings and an integer (input folder, variable
// folder, variable name, time step), each conversion checked before use
std::string format_file_name(const std::string &f, const std::string &a, const std::string &b,
                             const std::string &c, int t) {
  std::string out;
  int arg = 0;
  for (size_t i = 0; i < f.size(); ++i) {
    if (f[i] != '%') { out += f[i]; continue; }
    if (i + 1 < f.size() && f[i + 1] == '%') { out += '%'; ++i; continue; }
    size_t j = i + 1;
    while (j < f.size() && std::strchr("-+ #0123456789.", f[j])) ++j;
    if (j >= f.size()) throw Failure("fileFormatString: incomplete conversion");
    const std::string spec = f.substr(i, j - i + 1);
    const char conv = f[j];
    if (arg < 3 && conv == 's') {
      char buf[1024];
      std::snprintf(buf, sizeof buf, spec.c_str(), (arg == 0 ? a : arg == 1 ? b : c).c_str());
      out += buf;
    } else if (arg == 3 && (conv == 'd' || conv == 'i')) {
      char buf[64];
      std::snprintf(buf, sizeof buf, spec.c_str(), t);
      out += buf;
    } else {
      throw Failure("fileFormatString: expected %s, %s, %s, %d in that order");
    }
    ++arg;
    i = j;
  }
  return out;
}

int run_script(Gpu &gpu, const jl::Value &input) {
  for (const jl::Value &e : input.at("entries").arr) {
    const jl::Value &d = e.at("dimensions");
    if (d.arr.size() != 3) throw Failure("dimensions must have 3 entries");
    for (const jl::Value &var : e.at("variables").arr) {
      for (const jl::Value &tv : e.at("timesteps").arr) {
        const int t = (int)tv.as_int();
        SingleParams p;
        p.file_name = format_file_name(e.at("fileFormatString").as_str(),
                                       e.at("inputFolder").as_str(),
                                       var.at("variableFolder").as_str(), var.at("name").as_str(), t);
        std::printf("%s\n", p.file_name.c_str());
        if (!fs::exists(p.file_name + ".dat")) {
          std::printf("%s.dat\n", p.file_name.c_str());
          continue;
        }
        p.X = (int)d.arr[0].as_int();
        p.Y = (int)d.arr[1].as_int();
        p.Z = (int)d.arr[2].as_int();
        p.min_value = (float)var.at("minValue").as_num();
        p.max_value = (float)var.at("maxValue").as_num();
        p.output_folder = e.at("outputFolder").as_str();
        p.variable_name = var.at("name").as_str();
        p.timestep = t;
        p.json_file = e.at("jsonFile").as_str();
        p.ds = (int)e.at("dfDownscale").as_int();
        p.approximation = false;  // APPROXIMATION in the scripting build
        run_single(gpu, p);
      }
    }
  }
  return 0;
}

// ---- multi ----------------------------------------------------------------
// MultiVar::computeMultiMap (MultiVar.cpp:168-249): the 16 representative
// fields of each volume, every (i, 16+j) pair, H = HIST_SIZE
std::string multi_map(Gpu &gpu, const std::string &root, int ds, int X, int Y, int Z,
                      const std::string &v1, const std::string &v2, int t1, int t2,
                      const std::string &folder) {
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t F = (int64_t)(X / ds) * (Y / ds) * (Z / ds);
  const std::string p1 = root + "/" + v1 + "-" + std::to_string(t1) + "/distancefields/";
  const std::string p2 = root + "/" + v2 + "-" + std::to_string(t2) + "/distancefields/";
  if (!fs::exists(p1) || !fs::exists(p2)) return "failed";
  iso_ctx *ctx = gpu.get();
  DeviceBuf d_fields(gpu, sizeof(float) * (size_t)F * 2 * kReps);
  for (int s = 0; s < 2; ++s)
    for (int i = 0; i < kReps; ++i) {
      const std::string path = (s ? p2 : p1) + "distancefield-" + std::to_string(i) + ".df";
      if (!fs::exists(path)) {
        std::fprintf(stderr, "missing %s\n", path.c_str());
        return "failed";
      }
      const std::vector<float> f = read_floats(path, (size_t)F);
      gpu.check(iso_copy_to_device(ctx, d_fields.as<float>() + (size_t)(s * kReps + i) * F,
                                   f.data(), (size_t)F * 4),
                "copy field");
    }
  DeviceBuf d_mm(gpu, sizeof(float) * 4 * kReps);
  DeviceBuf d_map(gpu, sizeof(float) * kReps * kReps);
  gpu.check(iso_field_interval(ctx, d_fields.as<float>(), 2 * kReps, F, d_mm.as<float>(),
                               d_mm.as<float>() + 2 * kReps),
            "iso_field_interval");
  gpu.check(iso_hist_nmi(ctx, d_fields.as<float>(), kReps, F, kHistSize, 1,
                         ISO_COMPAT_HIST_CHUNKS | ISO_COMPAT_NMI32, d_mm.as<float>(),
                         d_mm.as<float>() + 2 * kReps,
                         d_map.as<float>(), nullptr, nullptr, nullptr),
            "iso_hist_nmi");
  std::vector<float> map(kReps * kReps);
  gpu.check(iso_copy_to_host(ctx, map.data(), d_map.p, map.size() * 4), "copy map");
  const std::string base = fmt("%s/TMP-multiMap-%d-%s-%d-%s-%d", folder.c_str(), kReps, v1.c_str(),
                               t1, v2.c_str(), t2);
  write_bytes(base + ".dat", map.data(), map.size() * 4);
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const std::string line = base + "," + std::to_string(secs) + "\n";
  write_bytes(root + "-multi-tmp/runtimes-tmp.csv", line.data(), line.size(), true);
  return base + ".dat";
}

int run_multi(Gpu &gpu, const jl::Value &cfg) {
  const std::string root = cfg.at("dataRoot").as_str();
  const std::string data_json = root + "/" + cfg.at("jsonFile").as_str();
  const int ds = (int)cfg.at("dfDownscale").as_int();
  std::vector<int> steps;
  for (const jl::Value &t : cfg.at("timeSteps").arr) steps.push_back((int)t.as_int());
  if (!fs::exists(data_json)) throw Failure("run information not found: " + data_json);
  jl::Value run = jl::parse(read_text(data_json));
  const jl::Value &dims = run.at("dimensions");
  if (dims.arr.size() != 3) throw Failure("dimensions must have 3 entries");
  const int X = (int)dims.arr[0].as_int(), Y = (int)dims.arr[1].as_int(),
            Z = (int)dims.arr[2].as_int();
  std::vector<std::string> vars;
  for (const jl::Value &v : run.at("variables").arr) vars.push_back(v.as_str());
  // maps already in the file are not computed again (createHostSetStr)
  std::set<std::tuple<std::string, std::string, int, int>> done;
  for (const jl::Value &m : run.at("similarityMaps").arr) {
    const long long a = m.at("var1").as_int(), b = m.at("var2").as_int();
    if (a < 0 || b < 0 || a >= (long long)vars.size() || b >= (long long)vars.size()) continue;
    done.insert({vars[(size_t)a], vars[(size_t)b], (int)m.at("t1").as_int(),
                 (int)m.at("t2").as_int()});
  }
  const std::string prefix = root + "-multi-tmp/";
  auto save = [&]() {
    const std::string text = jl::dump(run);
    write_bytes(data_json, text.data(), text.size());
  };
  auto one = [&](const std::string &v1, const std::string &v2, int t1, int t2,
                 const std::string &folder) {
    if (done.count({v1, v2, t1, t2})) return;
    const std::string file = multi_map(gpu, root, ds, X, Y, Z, v1, v2, t1, t2, folder);
    if (file != "failed") {
      add_similarity_map(run, v1, v2, t1, t2, file);
      done.insert({v1, v2, t1, t2});
    }
    std::printf("%s\n", file.c_str());
  };
  const int nv = (int)vars.size(), nt = (int)steps.size();
  for (int i = 0; i < nv; ++i)
    for (int j = i; j < nv; ++j)
      for (int k = 0; k < nt; ++k) {
        const int t1 = steps[(size_t)k];
        const std::string folder = prefix + vars[(size_t)i] + "-" + std::to_string(t1);
        fs::create_directories(folder);
        if (i != j) {
          one(vars[(size_t)i], vars[(size_t)j], t1, t1, folder);  // VSM
        } else {
          for (int l = k + 1; l < nt; ++l)
            one(vars[(size_t)i], vars[(size_t)j], t1, steps[(size_t)l], folder);  // TSM
          save();
        }
      }
  save();
  return 0;
}

// ---- distance --------------------------------------------------------------
// DistanceEqualization (runtime/src/DistanceEqualization.cpp) as the reference
// builds it (runtime/hdr/DistanceEqualization.h:19-53): NUM_DISTANCES 16
// (17 isovalues), no estimation iterations (MIN/MAX_ITERATIONS 0), then
// BIN_SEARCH_MAX_ITR + 1 = 181 refinement steps that move every other
// isovalue to where the two distance lines cross, switching the moving parity
// every BIN_SEARCH_STEP_ITR steps; jump handling, spikes, restarts, dense
// sampling and surface-to-surface distances are compiled out there and
// absent here.  Every step is one ISM of the 17 isosurfaces (voxel
// approximation, H = 17 bins) on the GPU.
constexpr int kNumDist = 16;              // NUM_DISTANCES
constexpr int kBinSearchMaxItr = 180;     // BIN_SEARCH_MAX_ITR
constexpr int kBinSearchStepItr = 5;      // BIN_SEARCH_STEP_ITR
constexpr float kJumpThresholdDiff = 0.5f;  // JUMP_THRESHOLD_DIFF
constexpr double kDampening = 0.66;       // DAMPENING_FACTOR (names the run entry)
// the starting isovalues: the reference's member initialiser
// (DistanceEqualization.h:221), which replaces the computed ones
// (DistanceEqualization.cpp:678-681)
constexpr float kStartIsovalues[kNumDist + 1] = {
    137.58087158203125f, 137.7071075439453125f, 138.2776336669921875f, 140.0638885498046875f,
    144.370941162109375f, 151.1020965576171875f, 157.6292877197265625f, 163.8123931884765625f,
    174.6049346923828125f, 214.08880615234375f, 324.660186767578125f, 604.75189208984375f,
    1200.432861328125f, 2765.340087890625f, 6707.39599609375f, 14770.724609375f,
    33156.9921875f};

struct DistanceModule {
  struct Bound {  // boundInformation
    float value = 0, d_left = 0, d_right = 0;
  };
  struct Iter {  // iterationInfo
    float iso[kNumDist + 1] = {};
    Bound lower[kNumDist + 1], upper[kNumDist + 1];
    float dist[kNumDist] = {};
    bool fixed_even = false;
    int iteration = 0;
  };
  struct Moving {  // movingValueAndDistance
    float isovalue = 0, d_right = 0;
  };

  Gpu &gpu;
  const SingleParams &p;
  iso_ctx *ctx = nullptr;
  float *d_img = nullptr;
  float real_min = 0, real_max = 0;
  int X, Y, Z, ds;
  std::string out, var_folder, sim_folder, json_path;
  float distances[kNumDist] = {};
  bool fixed[kNumDist + 1] = {}, pushed[kNumDist + 1] = {};
  int iterations = 0;
  float best_avg_error = 10e20f, best_max_error = 10e20f;
  float best_iso[kNumDist + 1];
  Moving best_moving[kNumDist];
  std::vector<Iter> infos;
  int max_itr = kBinSearchMaxItr;
  int ism_calls = 0;

  DistanceModule(Gpu &g, const SingleParams &params) : gpu(g), p(params) {
    X = p.X; Y = p.Y; Z = p.Z; ds = p.ds;
    std::copy(kStartIsovalues, kStartIsovalues + kNumDist + 1, best_iso);
    // init() + createFolder() (DistanceEqualization.cpp:954-1072); the output
    // folder is used as given (no separator added, unlike Single)
    out = p.output_folder;
    var_folder = out + p.variable_name + "-" + std::to_string(p.timestep);
    const std::string appendix =
        std::to_string(kNumDist + 1) + "-" + "-approx" + "-scale_" + std::to_string(kScale);
    sim_folder = var_folder + "/similarityMap-" + appendix + "/";
    for (const std::string &d : {var_folder + "/isosurfaces-" + appendix + "/", sim_folder,
                                 var_folder + "/distancefields/"})
      fs::create_directories(d);
    json_path = out + p.json_file;
    std::ofstream o1(var_folder + "/distances.csv", std::ios::app),
        o2(var_folder + "/isovalues.csv", std::ios::app),
        o3(var_folder + "/distances-split.csv", std::ios::app);
    o1 << "iterations" << ",";
    o2 << "iterations" << ",";
    for (int i = 0; i < kNumDist + 1; ++i) {
      if (i != kNumDist) o1 << "d" << i << ",";
      o2 << "iv" << i << ",";
    }
    o1 << "avgD,maxError,averageError,stdDev,normMaxE,normAvgE";
    o3 << "splits:" << std::endl;
    o1 << std::endl;
    o2 << std::endl;
    fixed[0] = fixed[kNumDist] = true;
    if (const char *e = std::getenv("ISO_DISTANCE_MAX_ITR")) max_itr = std::max(0, std::atoi(e));
  }

  // DistanceEqualization::checkIfAlreadyProcessed: any volume of this time step
  bool already_processed() const {
    if (!fs::exists(json_path)) return false;
    const jl::Value run = jl::parse(read_text(json_path));
    for (const jl::Value &v : run.at("volumes").arr)
      if (v.at("timestep").as_int() == p.timestep) return true;
    return false;
  }

  int64_t cells(float iso) {
    int64_t n = 0;
    const int st = iso_approx_extract(ctx, d_img, X, Y, Z, iso, nullptr, 0, &n);
    if (st != ISO_OK && st != ISO_ERR_CAPACITY) gpu.check(st, "iso_approx_extract");
    return n;
  }

  // approximation() (DistanceEqualization.cpp:84-186): an empty isosurface's
  // isovalue walks towards the next one in sixteenths until it has cells
  void approximation(float *iso) {
    for (int d = 0; d < kNumDist + 1; ++d) {
      float cur = iso[d];
      int64_t n = cells(cur);
      const float val = d < kNumDist ? iso[d + 1] : cur + (real_max - real_min) / kNumDist;
      float incr = (val - cur) / kNumDist;
      for (int guard = 0; n == 0; ++guard) {
        if (guard > 1000000) throw Failure("distance: no non-empty isosurface near isovalue " + std::to_string(d));
        cur += incr;
        if (cur > real_max || cur != cur) {
          // the reference reads isovalues[d-1] here, before the array for d = 0
          if (d == 0) throw Failure("distance: the first isovalue passed the volume's maximum");
          cur = iso[d - 1];
          incr = (iso[d] - cur) / 16;
          cur += incr;
        }
        if (d < kNumDist && cur > iso[d + 1]) {
          iso[d] = iso[d + 1];
          iso[d + 1] = cur;
          cur = iso[d];
        }
        iso[d] = cur;
        n = cells(cur);
      }
    }
  }

  // computeDistances (DistanceEqualization.cpp:560-624): d_j = 1 - NMI(j, j+1)
  // of the 17-field map (calculate_histogram_W with H = 17)
  void compute_distances(const float *iso, DeviceBuf &d_map) {
    constexpr int k = kNumDist + 1;
    std::vector<int> ids(k);
    iso_ism_stats st;
    gpu.check(iso_ism_volume(ctx, d_img, X, Y, Z, iso, k, ISO_EXTRACT_APPROX, ds, k, ISO_DF_LBVH,
                             ISO_COMPAT_ALL, d_map.as<float>(), nullptr, ids.data(), &st),
              "iso_ism_volume");
    ++ism_calls;
    if (st.n_nonempty != k) throw Failure("distance: an isosurface is empty after the search");
    std::vector<float> map((size_t)k * k);
    gpu.check(iso_copy_to_host(ctx, map.data(), d_map.p, map.size() * 4), "copy map");
    for (int j = 0; j < kNumDist; ++j) distances[j] = 1 - map[(size_t)j * k + j + 1];
  }

  static void report_array(std::ofstream &o, const float *a, int n) {
    for (int i = 0; i < n; ++i) o << std::setprecision(32) << a[i] << ",";
  }

  void report_isovalues(int iteration, bool final_run, const float *iso) {
    std::ofstream o(var_folder + "/isovalues.csv", std::ios::app);
    o << (final_run ? std::string("final") : std::to_string(iteration)) << "," << "";
    report_array(o, iso, kNumDist + 1);
    o << std::endl;
  }

  // reportDistances (DistanceEqualization.cpp:1637-1725)
  void report_distances(int iteration, bool final_run, float avg, float max_err, float avg_err,
                        const float *d, float std_dev, float n_max, float n_avg, bool jumps,
                        const float *jump, float std_nj) {
    std::ofstream o(var_folder + (jumps ? "/jumps.csv" : "/distances.csv"), std::ios::app);
    o << (final_run ? std::string("final") : std::to_string(iteration)) << ",";
    if (jumps && iteration == 0) {
      o << "iterations,";
      for (int i = 0; i < kNumDist; ++i) o << "d" << std::to_string(i) << ",";
      for (int i = 0; i < kNumDist; ++i) o << "j" << std::to_string(i) << ",";
      o << "stdNJ,avgD,maxError,averageError,stdDev,normMaxE,normAvgE";
      o << std::endl;
    }
    report_array(o, d, kNumDist);  // leaves the stream at 32 digits
    if (jumps) {
      report_array(o, jump, kNumDist);
      o << std_nj << ",";
    }
    o << avg << "," << max_err << "," << avg_err << "," << std_dev << "," << n_max << ","
      << n_avg;
    o << std::endl;
    std::ofstream o2(var_folder + "/distances-split.csv", std::ios::app);
    o2 << iteration << ",";
    std::vector<int> ids;
    for (int j = 0; j < kNumDist + 1; ++j)
      if (fixed[j]) ids.push_back(j);
    for (size_t j = 0; j + 1 < ids.size(); ++j) {
      float acc = 0.0f;
      for (int i = ids[j]; i < ids[j + 1]; ++i) acc += d[i];
      const float avg_dis = acc / (ids[j + 1] - ids[j]);
      o2 << avg_dis << ",";
    }
    o2 << std::endl;
  }

  // evaluateDistances (DistanceEqualization.cpp:333-558) with jump handling
  // off (HANDLE_JUMPS false): statistics, best tracking, moving-value bests
  bool evaluate(const float *iso, bool final_run, bool fix_even) {
    float acc = 0;
    for (int j = 0; j < kNumDist; ++j) acc += distances[j];
    const float avg = acc / kNumDist;
    float errors[kNumDist], acc_v = 0, max_err = -1;
    acc = 0;
    for (int k = 0; k < kNumDist; ++k) {
      errors[k] = std::abs(distances[k] - avg);
      acc_v += errors[k] * errors[k];
      acc += errors[k];
      if (errors[k] > max_err) max_err = errors[k];
    }
    const float avg_err = acc / kNumDist, variance = acc_v / kNumDist;
    const float std_dev = std::sqrt(variance);
    const float n_max = max_err / avg, n_avg = avg_err / avg;
    report_distances(iterations, final_run, avg, max_err, avg_err, distances, std_dev, n_max,
                     n_avg, false, nullptr, 0);
    const bool done = avg_err / avg > best_avg_error && max_err > best_max_error;
    if (avg_err / avg < best_avg_error) {
      best_avg_error = avg_err / avg;
      std::copy(iso, iso + kNumDist + 1, best_iso);
    }
    if (max_err < best_max_error) best_max_error = max_err;
    // jump statistics (computeJumps is forced on)
    float acc_nj = 0;
    int count = 0;
    float jump[kNumDist] = {};
    for (int i = fix_even ? 1 : 0; i < kNumDist - 1; i += 2) {
      if (i > 0) {  // TAKE_BEST_AT_FIX_SWITCH
        const float dl = distances[i - 1], dr = distances[i];
        const float diff = std::abs(dl - dr);
        const float best_diff = std::abs(best_moving[i - 1].d_right - best_moving[i].d_right);
        if (diff < best_diff) {
          best_moving[i - 1] = Moving{iso[i - 1], dl};
          best_moving[i] = Moving{iso[i], dr};
        }
      }
      float others = 0.0f;
      for (int j = 0; j < kNumDist; ++j)
        if (j != i && j != i + 1) others += distances[j];
      others /= (kNumDist - 2);
      const float err = std::abs(distances[i] - others) / others;
      jump[i + 1] = err;
      if (!(err > kJumpThresholdDiff)) {
        const float e1 = distances[i] - avg, e2 = distances[i + 1] - avg;
        acc_nj += e1 * e1 + e2 * e2;
        count += 2;
      }
    }
    acc_nj /= count;
    report_distances(iterations, final_run, avg, max_err, avg_err, distances, std_dev, n_max,
                     n_avg, true, jump, std::sqrt(acc_nj));
    return done;
  }

  // fixInversions (DistanceEqualization.cpp:192-215)
  void fix_inversions(float *iso) {
    for (int i = 0; i < kNumDist + 1; ++i)
      for (int j = i + 1; j < kNumDist + 1; ++j)
        if (iso[i] > iso[j]) {
          std::swap(iso[i], iso[j]);
          std::swap(fixed[i], fixed[j]);
          std::swap(pushed[i], pushed[j]);
        }
  }

  // binaryIsoValueSearch (DistanceEqualization.cpp:1338-1449): each moving
  // isovalue goes to the crossing of the lines through its bounds' distances
  void binary_iso_search(bool fixed_even, float *iso, Iter &it) {
    const int info_no = iterations % kBinSearchStepItr;
    for (int i = fixed_even ? 1 : 2; i < kNumDist; i += 2) {
      const int j = i - 1, k = i + 1;
      const float d_ji = distances[j], d_ik = distances[i];
      const float sum = d_ji + d_ik;
      Bound ub, lb;
      if (info_no == 0) {
        ub = Bound{iso[k], sum, 0.0f};
        lb = Bound{iso[j], 0.0f, sum};
      } else {
        ub = infos[(size_t)iterations - 1].upper[i];
        lb = infos[(size_t)iterations - 1].lower[i];
      }
      if (d_ji > d_ik) ub = Bound{iso[i], d_ji, d_ik};
      if (d_ji < d_ik) lb = Bound{iso[i], d_ji, d_ik};
      const float incr1 = (lb.d_left - ub.d_left) / (lb.value - ub.value);
      const float fix1 = lb.d_left - incr1 * lb.value;
      const float incr2 = (lb.d_right - ub.d_right) / (lb.value - ub.value);
      const float fix2 = lb.d_right - incr2 * lb.value;
      const float next = (fix2 - fix1) / (incr1 - incr2);
      if (!std::isnan(next) && !fixed[i]) iso[i] = next;
      it.upper[i] = ub;
      it.lower[i] = lb;
    }
  }

  // saveIteration + saveRunInfo (DistanceEqualization.cpp:776-793, 1282-1310)
  void save_iteration(const float *iso, int iteration) {
    jl::Value run = fs::exists(json_path) ? jl::parse(read_text(json_path))
                                          : new_run_info(X, Y, Z, ds);
    const std::string name = p.variable_name + "-" + std::to_string(kDampening);
    jl::Value vol = jl::Value::object();
    vol.set("variable", jl::Value::of_int(var_index(run, name)));
    vol.set("timestep", jl::Value::of_int(p.timestep));
    vol.set("numIsovalues", jl::Value::of_int(kNumDist + 1));
    jl::Value ra = jl::Value::array();
    for (int i = 0; i < kNumDist + 1; ++i) {
      jl::Value o = jl::Value::object();
      o.set("valueId", jl::Value::of_int(i));
      o.set("isovalue", jl::Value::of_float(iso[i]));
      o.set("repId", jl::Value::of_int(i));
      o.set("importance", jl::Value::of_float((float)i));
      o.set("filename", jl::Value::of(std::to_string(iteration) + "-surface-" +
                                      std::to_string(i) + ".dat"));
      o.set("mapId", jl::Value::of_int(i));
      ra.arr.push_back(o);
    }
    vol.set("representatives", ra);
    vol.set("runtime", jl::Value::of_float((double)iteration));
    run.at("volumes").arr.push_back(vol);
    add_similarity_map(run, name, name, p.timestep, p.timestep,
                       sim_folder + "map-" + std::to_string(kDampening) + ".dat");
    const std::string text = jl::dump(run);
    write_bytes(json_path, text.data(), text.size());
  }

  // binarySearchEqualization (DistanceEqualization.cpp:1451-1605), its tail
  // recursion as a loop.  A step starts from the previous step's isovalues;
  // at a parity switch the distances (not the isovalues) are replaced by the
  // best moving values, as in the reference.
  void refine(DeviceBuf &d_map) {
    bool fixed_even = true, start = true;
    float iso[kNumDist + 1];
    while (true) {
      if (iterations == 0) {
        fixed_even = true;
        if (start) {
          std::copy(best_iso, best_iso + kNumDist + 1, iso);
          for (int j = 0; j < kNumDist; ++j) best_moving[j] = Moving{best_iso[j], distances[j]};
        }
      } else {
        std::copy(infos.back().iso, infos.back().iso + kNumDist + 1, iso);
      }
      Iter it;
      fix_inversions(iso);
      binary_iso_search(fixed_even, iso, it);
      approximation(iso);
      compute_distances(iso, d_map);
      const bool finished = iterations == max_itr;
      report_isovalues(iterations, finished, iso);
      evaluate(iso, finished, fixed_even);
      it.iteration = iterations;
      it.fixed_even = fixed_even;
      std::copy(iso, iso + kNumDist + 1, it.iso);
      std::copy(distances, distances + kNumDist, it.dist);
      infos.push_back(it);
      if (iterations % 10 == 0) save_iteration(iso, iterations);
      if (finished) {
        save_iteration(iso, iterations);
        return;
      }
      ++iterations;
      if (iterations % kBinSearchStepItr == 0) {
        for (int i = 0; i < kNumDist; ++i) distances[i] = best_moving[i].d_right;
        fixed_even = !fixed_even;
      }
      start = false;
    }
  }

  // DistanceEqualization::run(nullptr, false, false) (DistanceEqualization.cpp:626-773)
  int run() {
    if (already_processed()) {
      std::printf("timestep %d already processed... skipping\n", p.timestep);
      return 0;
    }
    const auto t0 = std::chrono::steady_clock::now();
    if (X < 2 || Y < 2 || Z < 2 || ds < 1 || X < ds || Y < ds || Z < ds)
      throw Failure("bad dimensions or dfDownscale");
    const size_t N = (size_t)X * Y * Z;
    std::vector<float> file = read_floats(p.file_name + ".dat", N);
    ctx = gpu.get();
    DeviceBuf img(gpu, N * 4);
    d_img = img.as<float>();
    {
      DeviceBuf d_file(gpu, N * 4);
      gpu.check(iso_copy_to_device(ctx, d_file.p, file.data(), N * 4), "copy");
      gpu.check(iso_volume_ingest(ctx, d_file.as<float>(), X, Y, Z, p.min_value, p.max_value,
                                  d_img, &real_min, &real_max),
                "iso_volume_ingest");
    }
    file.clear();
    file.shrink_to_fit();
    DeviceBuf d_map(gpu, sizeof(float) * (kNumDist + 1) * (kNumDist + 1));
    ++iterations;
    // the computed (equalised) isovalues are all replaced by the start table
    float iso[kNumDist + 1];
    std::copy(best_iso, best_iso + kNumDist + 1, iso);
    approximation(iso);
    compute_distances(iso, d_map);
    evaluate(iso, false, false);
    // MIN_ITERATIONS == MAX_ITERATIONS == 0: this is the final estimation
    report_isovalues(iterations, true, iso);
    iterations = 0;
    refine(d_map);
    std::copy(iso, iso + kNumDist + 1, best_iso);
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("distance %s t=%d: %d ISMs of %d isosurfaces in %.3f s, best normalised average "
                "error %g\n",
                p.variable_name.c_str(), p.timestep, ism_calls, kNumDist + 1, secs,
                (double)best_avg_error);
    return 0;
  }
};

int usage() {
  std::fprintf(stderr,
               "usage: isosurfaces_amd <module> <config.json>\n"
               "  single\t computes a single ISM (voxel approximation)\n"
               "  script\t computes the ISMs of a scripting input (marching cubes)\n"
               "  multi \t computes TSM/VSM maps of saved representative fields\n"
               "  distance\t equalises the NMI distances of 17 isosurfaces (voxel approximation)\n");
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) return usage();
  const std::string module = argv[1];
  try {
    Gpu gpu;
    const jl::Value cfg = jl::parse(read_text(argv[2]));
    if (module == "single") return run_single(gpu, single_from_config(cfg));
    if (module == "script") return run_script(gpu, cfg);
    if (module == "multi") return run_multi(gpu, cfg);
    if (module == "distance") {
      const SingleParams p = single_from_config(cfg);
      DistanceModule dm(gpu, p);
      return dm.run();
    }
    return usage();
  } catch (const std::exception &e) {
    std::fprintf(stderr, "isosurfaces_amd %s: %s\n", module.c_str(), e.what());
    return 2;
  }
}
