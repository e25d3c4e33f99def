// mlem -- MLEM/OSEM reconstruction driver over the C ABI (src/bin/mlem.rs:47-105).
//
//   mlem <config.toml> <output-directory>
//        [--lors FILE | --records FILE --discretize "R_MIN DR DZ DA" | --synthetic N [--tof-fwhm PS]] [--seed S]
//        [--autotune]
//
// Reads the config, copies it to <output-directory>/mlem-config.toml, loads
// the sensitivity image named in [attenuation_correction] (sizes must match
// the FOV exactly), runs iterations.number * iterations.subsets subset
// updates and writes each image as "<output-directory><it:02>-<sub:02>.raw"
// (Image3D).  As in the reference, the directory string is concatenated
// without a separator.  LORs come from a file of 32-byte records
// (petalo_lor_t), from the seeded synthetic phantom, or from the raw bytes of
// the reco_info/lors dataset (44-byte Hdf5Lor records, e.g. written by h5py's
// `f["reco_info/lors"][:].tofile(path)`): these go through the config's cuts,
// energy smearing, RandomZPhi position smearing with the dataset's
// discretization (--discretize, its r_min/dr/dz/da attributes in mm) and
// scatter correction, as io::hdf5::read_lors does (src/io/hdf5.rs:186-228);
// the HDF5 file read itself is not part of this library.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "petalo_hip.h"

namespace {

[[noreturn]] void die(const std::string& m) {
  std::fprintf(stderr, "mlem: %s\n", m.c_str());
  std::exit(1);
}

void check(int rc, const char* what) {
  if (rc) die(std::string(what) + ": " + petalo_last_error());
}

struct Progress {  // utils::timing::Progress (src/utils.rs:94-138)
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void done(const std::string& msg) {
    const auto now = std::chrono::steady_clock::now();
    std::printf("%s: %lld ms\n", msg.c_str(),
                (long long)std::chrono::duration_cast<std::chrono::milliseconds>(now - t).count());
    std::fflush(stdout);
    t = now;
  }
};

void mkdir_p(const std::string& path) {
  std::string cur;
  for (size_t i = 0; i < path.size(); ++i) {
    cur += path[i];
    if (path[i] == '/' || i + 1 == path.size()) mkdir(cur.c_str(), 0755);
  }
}

// Float equality within 1 ulp (float_eq! ulps_all <= 1, mlem.rs:112-130).
bool ulps1(float a, float b) {
  int32_t ia, ib;
  std::memcpy(&ia, &a, 4);
  std::memcpy(&ib, &b, 4);
  return std::abs((long long)ia - (long long)ib) <= 1;
}

}  // namespace

int main(int argc, char** argv) {
  const bool help = argc >= 2 && (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help"));
  if (argc < 3 || help) {
    std::printf("Maximum Likelyhood Expectation Maximization\n\n"
                "usage: mlem <config.toml> <output-directory> [--lors FILE | --records FILE --discretize "
                "\"R_MIN DR DZ DA\" | --synthetic N [--tof-fwhm PS]] [--seed S] [--autotune]\n");
    return help ? 0 : 1;
  }
  const std::string cfg_path = argv[1], outdir = argv[2];
  std::string lor_file, record_file, discretize;
  size_t synth_n = 0;
  uint64_t seed = 1;
  float fwhm = 0.0f;
  bool tune = false;
  for (int i = 3; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) die("missing value for " + a);
      return argv[++i];
    };
    if (a == "--lors") lor_file = val();
    else if (a == "--records") record_file = val();
    else if (a == "--discretize") discretize = val();
    else if (a == "--synthetic") synth_n = std::strtoull(val().c_str(), nullptr, 10);
    else if (a == "--seed") seed = std::strtoull(val().c_str(), nullptr, 10);
    else if (a == "--tof-fwhm") fwhm = std::strtof(val().c_str(), nullptr);
    else if (a == "--autotune") tune = true;
    else if (a == "-j" || a == "--mlem-threads") (void)val();  // accepted for compatibility; the GPU ignores it
    else die("unknown argument " + a);
  }

  Progress progress;
  petalo_config_t cfg;
  check(petalo_config_read(cfg_path.c_str(), &cfg), "reading config");
  mkdir_p(outdir);
  {
    std::ifstream src(cfg_path, std::ios::binary);
    std::ofstream dst(outdir + "/mlem-config.toml", std::ios::binary);
    dst << src.rdbuf();
    if (!dst) die("Cannot write in output directory `" + outdir + "`");
  }
  std::printf("Configuration:\n[iterations] number = %llu, subsets = %llu\n[fov] nvoxels = [%u, %u, %u], "
              "size = [%g mm, %g mm, %g mm]\n",
              (unsigned long long)cfg.iterations_number, (unsigned long long)cfg.iterations_subsets, cfg.fov.n[0],
              cfg.fov.n[1], cfg.fov.n[2], cfg.fov.full_size_mm[0], cfg.fov.full_size_mm[1], cfg.fov.full_size_mm[2]);
  if (cfg.has_tof) std::printf("[tof] sigma = %g ps, cutoff = %g\n", cfg.tof.sigma_ps, cfg.tof.cutoff);
  progress.done("Startup");

  const size_t nv = (size_t)cfg.fov.n[0] * cfg.fov.n[1] * cfg.fov.n[2];
  std::vector<float> sens;
  if (cfg.has_attenuation_correction) {
    petalo_fov_t sf;
    check(petalo_image3d_read_header(cfg.sensitivity_image, &sf), "reading sensitivity image");
    float hw_a[3], vs_a[3], hw_b[3], vs_b[3];
    petalo_fov_derive(&sf, hw_a, vs_a);
    petalo_fov_derive(&cfg.fov, hw_b, vs_b);
    bool ok = true;
    for (int i = 0; i < 3; ++i)
      ok = ok && sf.n[i] == cfg.fov.n[i] && ulps1(hw_a[i] * 2.0f, hw_b[i] * 2.0f);
    if (!ok) die("For now, the sensitivity image must match the dimensions of the output image exactly.");
    sens.resize(nv);
    check(petalo_image3d_read(cfg.sensitivity_image, nullptr, sens.data(), nv), "reading sensitivity image");
    progress.done("Loaded sensitivity image");
  }

  std::vector<petalo_lor_t> lors;
  if (!lor_file.empty()) {
    std::ifstream f(lor_file, std::ios::binary | std::ios::ate);
    if (!f) die("cannot read " + lor_file);
    const size_t bytes = (size_t)f.tellg();
    if (bytes % sizeof(petalo_lor_t)) die("LOR file size is not a multiple of 32 bytes");
    lors.resize(bytes / sizeof(petalo_lor_t));
    f.seekg(0);
    f.read(reinterpret_cast<char*>(lors.data()), (std::streamsize)bytes);
  } else if (!record_file.empty()) {
    float d[4];
    if (std::sscanf(discretize.c_str(), "%f %f %f %f", d, d + 1, d + 2, d + 3) != 4)
      die("--records needs --discretize \"R_MIN DR DZ DA\" (mm)");
    std::ifstream f(record_file, std::ios::binary | std::ios::ate);
    if (!f) die("cannot read " + record_file);
    const size_t bytes = (size_t)f.tellg();
    if (bytes % 44) die("record file size is not a multiple of 44 bytes");
    std::vector<float> rec(bytes / 4);
    f.seekg(0);
    f.read(reinterpret_cast<char*>(rec.data()), (std::streamsize)bytes);
    petalo_ingest_t ip;
    check(petalo_config_ingest(&cfg, d[0], d[1], d[2], d[3], seed, &ip), "ingestion parameters");
    const size_t n_rec = bytes / 44;
    lors.resize(n_rec);
    size_t n_out = 0;
    uint64_t n_read = 0;
    check(petalo_ingest_lors(rec.data(), n_rec, &ip, cfg.has_scatter_correction ? &cfg.scatter : nullptr,
                             lors.data(), lors.size(), &n_out, &n_read),
          "ingesting LORs");
    lors.resize(n_out);
    std::printf("   Reading LORs: %llu read, %zu selected\n", (unsigned long long)n_read, n_out);
  } else if (synth_n) {
    lors.resize(synth_n);
    check(petalo_synth_phantom_lors(seed, synth_n, fwhm, lors.data()), "synthetic LORs");
  } else {
    die("no LORs: give --lors FILE, --records FILE or --synthetic N");
  }
  progress.done("Loaded " + std::to_string(lors.size()) + " LORs");

  petalo_mlem_t* ctx = nullptr;
  petalo_mlem_options_t opt{PETALO_SCHEME_AUTO, 1, 0, {0, 0, 0}, tune ? 1 : 0, PETALO_PRECISION_F32};
  check(petalo_mlem_create(&ctx, &cfg.fov, cfg.has_tof ? &cfg.tof : nullptr, lors.data(), lors.size(),
                           sens.empty() ? nullptr : sens.data(), cfg.iterations_subsets, &opt),
        "creating MLEM context");
  std::vector<float> image(nv);
  const uint64_t total = cfg.iterations_number * cfg.iterations_subsets;
  for (uint64_t k = 0; k < total; ++k) {
    petalo_osem_t id;
    check(petalo_mlem_next(ctx, image.data(), &id), "MLEM update");
    char name[64];
    std::snprintf(name, sizeof name, "Iteration %2llu-%02llu", (unsigned long long)id.iteration,
                  (unsigned long long)id.subset);
    progress.done(name);
    char file[64];
    std::snprintf(file, sizeof file, "%02llu-%02llu.raw", (unsigned long long)id.iteration,
                  (unsigned long long)id.subset);
    check(petalo_image3d_write((outdir + file).c_str(), &cfg.fov, image.data()), "writing image");
    progress.done("                               Wrote raw bin");
  }
  petalo_mlem_destroy(ctx);
  return 0;
}
