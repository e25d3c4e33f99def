// foms -- NEMA7 or Jaszczak figures of merit of a raw image over the C ABI
// (src/bin/foms.rs).
//
//   foms nema7|jaszczak <image.raw>
//
// Prints, per hot sphere, its diameter, contrast recovery coefficient,
// background variability and signal-to-noise ratio (percentages), and for
// NEMA7 the accuracy of corrections in the lung insert per slice.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "petalo_hip.h"

namespace {

[[noreturn]] void die(const std::string& m) {
  std::fprintf(stderr, "foms: %s\n", m.c_str());
  std::exit(1);
}

void check(int rc, const char* what) {
  if (rc) die(std::string(what) + ": " + petalo_last_error());
}

}  // namespace

int main(int argc, char** argv) {
  const bool help = argc >= 2 && (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help"));
  if (argc != 3 || help) {
    std::printf("Calculate NEMA7 or Jaszczak Figures of Merit (FOMs) from raw image files\n\n"
                "usage: foms <nema7|jaszczak> <input-file>\n");
    return help ? 0 : 1;
  }
  const std::string which = argv[1];
  int32_t phantom;
  if (which == "nema7") phantom = PETALO_PHANTOM_NEMA7;
  else if (which == "jaszczak") phantom = PETALO_PHANTOM_JASZCZAK;
  else die("unknown phantom '" + which + "' (nema7, jaszczak)");
  petalo_fov_t fov;
  check(petalo_image3d_read_header(argv[2], &fov), "reading image");
  const size_t nv = (size_t)fov.n[0] * fov.n[1] * fov.n[2];
  std::vector<float> image(nv);
  check(petalo_image3d_read(argv[2], nullptr, image.data(), nv), "reading image");
  petalo_fom_t foms[6];
  std::vector<float> aocs(4096);
  size_t n_aocs = 0;
  check(petalo_phantom_foms(&fov, image.data(), phantom, foms, aocs.data(), aocs.size(), &n_aocs), "FOMs");
  std::printf("Sphere diameter / mm    CRC %%   bg var %%    SNR %%\n");
  for (const petalo_fom_t& f : foms)
    std::printf("%20.1f %8.1f %8.1f %10.1f\n", f.r_mm * 2.0f, f.crc, f.bg_variability, f.snr);
  if (phantom == PETALO_PHANTOM_NEMA7) {
    std::printf("\nAOCs: [");
    for (size_t i = 0; i < n_aocs && i < aocs.size(); ++i) std::printf(i ? ", %.1f" : "%.1f", aocs[i]);
    std::printf("]\n");
  }
  return 0;
}
