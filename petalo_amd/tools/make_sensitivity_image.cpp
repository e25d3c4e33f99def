// make_sensitivity_image -- sensitivity image from a density image over the
// C ABI (src/bin/make_sensitivity_image/main.rs:6-74, cli.rs:1-65).
//
//   make_sensitivity_image -i density.raw [-o out.raw] -l LENGTH -r RADIUS
//       [--rho-to-mu 0.095] [--output-image-dims '10 cm 20 cm 255 mm 30 30 35']
//       [--seed S] continuous N_LORS
//   make_sensitivity_image ... discrete --dr DR --dz DZ --da DA
//       --adjust no|centre|random|random-z-phi
//
// The density image [kg/m^3] becomes mu [1/mm]; N_LORS random LORs with both
// ends on the detector cylinder (length, inner radius) that enter the FOV
// are backprojected with weights exp(+integral of mu) and the image is
// normalised by N_LORS.  The LORs come from a seeded counter-based stream
// (the reference draws them from an unseeded generator).  The discrete
// model backprojects every pair of detector-element centres (elements of
// DR x DZ x DA at inner radius RADIUS), both ends adjusted inside their
// element, normalised by n(n-1)/2 over the n elements
// (src/bin/make_sensitivity_image/discrete.rs:3-62).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "petalo_hip.h"

namespace {

[[noreturn]] void die(const std::string& m) {
  std::fprintf(stderr, "make_sensitivity_image: %s\n", m.c_str());
  std::exit(1);
}

void check(int rc, const char* what) {
  if (rc) die(std::string(what) + ": " + petalo_last_error());
}

}  // namespace

int main(int argc, char** argv) {
  std::string in, out = "sensitivity.raw", dims;
  float L = 0, R = 0, rho_to_mu = 0.095f;
  uint64_t seed = 0;
  size_t n_lors = 0;
  bool have_type = false, discrete = false;
  float dr = 0, dz = 0, da = 0;
  int32_t adjust = -1;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) die("missing value for " + a);
      return argv[++i];
    };
    auto len = [&](const std::string& s) {
      float v;
      check(petalo_parse_quantity(s.c_str(), 'L', &v), "parsing length");
      return v;
    };
    if (a == "-h" || a == "--help") {
      std::printf("Create sensitivity image from a density image\n\nusage: make_sensitivity_image -i IN [-o OUT] "
                  "-l LENGTH -r RADIUS [--rho-to-mu X] [--output-image-dims DIMS] [--seed S]\n"
                  "       (continuous N | discrete --dr DR --dz DZ --da DA --adjust no|centre|random|random-z-phi)\n");
      return 0;
    }
    if (a == "-i" || a == "--input") in = val();
    else if (a == "-o" || a == "--output") out = val();
    else if (a == "-l" || a == "--detector-length") L = len(val());
    else if (a == "-r" || a == "--r-min") R = len(val());
    else if (a == "--rho-to-mu") rho_to_mu = std::strtof(val().c_str(), nullptr);
    else if (a == "--output-image-dims") dims = val();
    else if (a == "--seed") seed = std::strtoull(val().c_str(), nullptr, 10);
    else if (a == "-j" || a == "--n-threads") (void)val();
    else if (a == "continuous") {
      n_lors = std::strtoull(val().c_str(), nullptr, 10);
      have_type = true;
    } else if (a == "discrete") {
      have_type = discrete = true;
    } else if (discrete && a == "--dr") dr = len(val());
    else if (discrete && a == "--dz") dz = len(val());
    else if (discrete && a == "--da") da = len(val());
    else if (discrete && a == "--adjust") {
      // clap's value names of discrete::Adjust (src/discrete.rs:24-41)
      const std::string v = val();
      adjust = v == "no" ? PETALO_ADJUST_NONE
               : v == "centre" ? PETALO_ADJUST_CENTRE
               : v == "random" ? PETALO_ADJUST_RANDOM
               : v == "random-z-phi" ? PETALO_ADJUST_RANDOM_Z_PHI
                                     : -1;
      if (adjust < 0) die("unknown --adjust " + v + " (no, centre, random, random-z-phi)");
    } else
      die("unknown argument " + a);
  }
  if (in.empty() || !(L > 0) || !(R > 0) || !have_type) die("need -i, -l, -r and 'continuous N' or 'discrete ...'");
  if (discrete && (!(dr > 0) || !(dz > 0) || !(da > 0) || adjust < 0))
    die("discrete needs --dr, --dz, --da and --adjust");

  auto t0 = std::chrono::steady_clock::now();
  auto report = [&](const std::string& msg) {
    auto t = std::chrono::steady_clock::now();
    std::printf("%s: %lld ms\n", msg.c_str(),
                (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t - t0).count());
    std::fflush(stdout);
    t0 = t;
  };
  report("Startup");
  petalo_fov_t fov;
  check(petalo_image3d_read_header(in.c_str(), &fov), "reading density image");
  const size_t nv = (size_t)fov.n[0] * fov.n[1] * fov.n[2];
  std::vector<float> mu(nv);
  check(petalo_image3d_read(in.c_str(), nullptr, mu.data(), nv), "reading density image");
  report("Read density image \"" + in + "\"");
  check(petalo_density_to_attenuation(mu.data(), nv, rho_to_mu), "density to attenuation");

  petalo_fov_t ofov = fov;
  if (!dims.empty()) check(petalo_fov_parse(dims.c_str(), &ofov), "parsing --output-image-dims");
  const size_t no = (size_t)ofov.n[0] * ofov.n[1] * ofov.n[2];
  std::vector<float> sens(no);
  petalo_stats_t st;
  if (discrete) {
    check(petalo_make_sensitivity_image_discrete(&fov, mu.data(), dims.empty() ? nullptr : &ofov, L, R, dr, dz, da,
                                                 adjust, seed, sens.data(), &st),
          "sensitivity image");
    std::printf("%llu LORs between detector elements\n", (unsigned long long)st.n_lors);
  } else {
    std::printf("Creating sensitivity image, using %zu LORs ... ", n_lors);
    std::fflush(stdout);
    check(petalo_make_sensitivity_image(&fov, mu.data(), dims.empty() ? nullptr : &ofov, nullptr, n_lors, seed, L,
                                        R * 2.0f, sens.data(), &st),
          "sensitivity image");
  }
  report("done");
  check(petalo_image3d_write(out.c_str(), &ofov, sens.data()), "writing sensitivity image");
  report("Wrote sensitivity image to \"" + out + "\"");
  return 0;
}
