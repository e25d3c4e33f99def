/*
 * petalo_oracle.h -- CPU oracle for the petalo reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in petalo_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product never links or calls it.
 *
 * It restates, operation for operation in IEEE binary32 (and again in
 * binary64 as the "FP64 oracle"), the Rust reference:
 *   walk/setup      src/projectors/siddon.rs:81-340
 *   FOV             src/fov.rs:22-71   (ray/box entry per ncollide3d 0.33.0)
 *   norm/normalize  geometry/src/vector.rs:160-167 (nalgebra 0.30.1)
 *   TOF Gaussian    src/gauss.rs:31-60
 *   projection      src/projector.rs:33-153
 *   MLEM update     src/mlem.rs:33-66
 *   sensitivity     src/bin/make_sensitivity_image/continuous.rs:3-17,
 *                   main.rs:77-94
 *   FOMs            src/fom.rs:34-461, src/bin/foms.rs:42-250 (fom_oracle.cpp)
 * plus an independent classic Siddon (parametric alpha merge, Siddon 1985)
 * and a brute-force per-voxel clipper, both in binary64, used only to pin
 * the restatement against the mathematics.
 */
#ifndef PETALO_ORACLE_H
#define PETALO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 32-byte LOR record: endpoints in mm, dt = t2 - t1 in ps (src/lor.rs:11-25). */
typedef struct {
  float p1[3];
  float p2[3];
  float dt_ps;
  float additive_correction;
} orc_lor_t;

/* TOF parameters as in the config (src/config/mlem.rs:177-188). */
typedef struct {
  float sigma_ps;
  float cutoff;
} orc_tof_t;

/* Work counters for one projection pass. */
typedef struct {
  uint64_t n_lors;       /* LORs visited                                  */
  uint64_t n_steps;      /* iterations of the walk loop (siddon.rs:96-129) */
  uint64_t n_nz;         /* pushed row elements (weight > 0)              */
  uint64_t n_skipped;    /* LORs dropped by the index check (projector.rs:103-118) */
  uint64_t n_missed;     /* LORs that miss the FOV (empty row)            */
  uint64_t n_runaway;    /* walks stopped at the step cap (reference would not terminate) */
  uint64_t n_shifted;    /* LORs whose walk starts one voxel short of its entry (an entry
                            coordinate e in (-1, -1e-5]: floor saturates to voxel 0) */
} orc_stats_t;

/* FOV derivation, binary32 (src/fov.rs:22-39). */
void orc_fov_derive_f32(const float full_mm[3], const uint32_t n[3],
                        float half_width[3], float voxel_size[3]);

/* FOV::entry (src/fov.rs:55-71); returns 1 and the entry point on a hit. */
int orc_fov_entry_f32(const float full_mm[3], const uint32_t n[3],
                      const float p1[3], const float p2[3], float entry[3]);

/* TOF Gaussian g(dx) in 1/mm (src/gauss.rs:31-60). */
float orc_gauss_f32(float sigma_ps, float cutoff, float dx_mm);
double orc_gauss_f64(double sigma_ps, double cutoff, double dx_mm);

/* One system-matrix row (siddon.rs:50-130).  tof may be NULL.  Writes at most
 * cap elements; returns the number of pushed elements (may exceed cap, in
 * which case only cap were written).  *steps gets the walk-loop count. */
int64_t orc_row_f32(const float full_mm[3], const uint32_t n[3],
                    const orc_tof_t* tof, const orc_lor_t* lor,
                    int64_t* idx, float* w, int64_t cap, uint64_t* steps);
int64_t orc_row_f64(const float full_mm[3], const uint32_t n[3],
                    const orc_tof_t* tof, const orc_lor_t* lor,
                    int64_t* idx, double* w, int64_t cap, uint64_t* steps);

/* Rows and MLEM updates with binary32 walk decisions and binary64 arithmetic:
 * the binary32 walk fixes hit, first voxel, start fraction (with its entry
 * quirk), the crossed axis of every step and which steps push; weights,
 * projections and the image are then binary64 along that step sequence.
 * Against the binary32 oracle this differs by rounding only, against the
 * binary64 oracle by the walk's discrete choices only. */
int64_t orc_row_mixed(const float full_mm[3], const uint32_t n[3],
                      const orc_tof_t* tof, const orc_lor_t* lor,
                      int64_t* idx, double* w, int64_t cap, uint64_t* steps);
int orc_mlem_update_mixed(const float full_mm[3], const uint32_t n[3],
                          const orc_tof_t* tof, const orc_lor_t* lors, size_t n_lors,
                          const double* sensitivity, double* image, int n_threads,
                          orc_stats_t* stats);

/* The library's sensitivity-LOR stream restated (continuous.rs:20-47 with
 * the library's Philox draws): attempt k of LOR i takes block (i, k, 0,
 * 0x5e45); z = L (u - 1/2), theta = TAU u, r = D/2 for both ends; kept when
 * FOV::entry hits.  attempts[i] = draws used (0: none found in 4096).
 * Returns the number of LORs found. */
int64_t orc_sensitivity_lors(const float full_mm[3], const uint32_t n[3], size_t n_lors, uint64_t seed,
                             float detector_length_mm, float detector_diameter_mm, orc_lor_t* out,
                             uint32_t* attempts);

/* FOV entry arithmetic of every later call (process-wide; test use only):
 * 0 (default) the restated ncollide3d 0.33 slab cast, entry = p1 + dir*toi;
 * 1 the same with the hit face's coordinate set to exactly -half_width (an
 * exact cast: no walk starts shifted).  Returns -1 for an unknown variant. */
int orc_set_entry_variant(int variant);
/* 1 when the binary32 walk of this LOR starts shifted (an entry coordinate e
 * in (-1, -1e-5] after the EPS snap of siddon.rs:134-153, floor saturating to
 * 0 at siddon.rs:312-313), 0 when not, -1 on a miss. */
int orc_entry_shifted_f32(const float full_mm[3], const uint32_t n[3], const orc_lor_t* lor);

/* Independent classic Siddon in binary64: returns the number of segments
 * (voxel index in the reference's 1-D layout, and length in mm), in
 * traversal order p1 -> p2, zero-length segments dropped. */
int64_t orc_classic_siddon_f64(const double full_mm[3], const uint32_t n[3],
                               const double p1[3], const double p2[3],
                               int64_t* idx, double* len, int64_t cap);

/* Brute force: length of segment [p1,p2] inside every voxel (binary64),
 * written to len_per_voxel[N_vox]. */
void orc_bruteforce_lengths_f64(const double full_mm[3], const uint32_t n[3],
                                const double p1[3], const double p2[3],
                                double* len_per_voxel);

/* Chord length of the segment [p1,p2] inside the FOV box (binary64 slab). */
double orc_chord_f64(const double full_mm[3], const double p1[3], const double p2[3]);

/* One MLEM/OSEM subset update over lors[0..n_lors) (src/mlem.rs:33-52).
 * sensitivity may be NULL (ones).  image is updated in place.  Work is cut
 * into 10 000-LOR jobs spread over n_threads threads, each folding into a
 * private backprojection that is then summed in thread order
 * (src/mlem.rs:57-66, src/projector.rs:48-62). */
int orc_mlem_update_f32(const float full_mm[3], const uint32_t n[3],
                        const orc_tof_t* tof, const orc_lor_t* lors, size_t n_lors,
                        const float* sensitivity, float* image, int n_threads,
                        orc_stats_t* stats);
int orc_mlem_update_f64(const float full_mm[3], const uint32_t n[3],
                        const orc_tof_t* tof, const orc_lor_t* lors, size_t n_lors,
                        const double* sensitivity, double* image, int n_threads,
                        orc_stats_t* stats);

/* Backprojection only (no update): bp[j] = sum over LORs of w_ij / (p_i * ac_i)
 * with p_i the forward projection of image. */
int orc_backproject_mlem_f32(const float full_mm[3], const uint32_t n[3],
                             const orc_tof_t* tof, const orc_lor_t* lors, size_t n_lors,
                             const float* image, float* bp, int n_threads,
                             orc_stats_t* stats);

/* Sensitivity image (continuous.rs:3-17 with project_one_lor_sens,
 * projector.rs:72-74): forward walk on the attenuation FOV, backward walk on
 * out FOV (NULL = same), normalised by n_lors (main.rs:77). */
int orc_sensitivity_f32(const float att_full_mm[3], const uint32_t att_n[3],
                        const float* mu,
                        const float* out_full_mm, const uint32_t* out_n,
                        const orc_lor_t* lors, size_t n_lors, int n_threads,
                        float* out, orc_stats_t* stats);
int orc_sensitivity_f64(const float att_full_mm[3], const uint32_t att_n[3],
                        const float* mu,
                        const float* out_full_mm, const uint32_t* out_n,
                        const orc_lor_t* lors, size_t n_lors, int n_threads,
                        double* out, orc_stats_t* stats);

/* density [kg/m^3] -> mu [1/mm] factor, binary32, as
 * src/bin/make_sensitivity_image/main.rs:11-17,80-94 compute it. */
float orc_rho_to_mu_factor_f32(float rho_to_mu_cm2_per_g);

/* Walk statistics only (no image): full reference walk step counts. */
int orc_walk_stats_f32(const float full_mm[3], const uint32_t n[3],
                       const orc_tof_t* tof, const orc_lor_t* lors, size_t n_lors,
                       int n_threads, orc_stats_t* stats);

/* LOR ingestion (src/io/hdf5.rs:97-228, src/discrete.rs:60-213,
 * src/lorogram.rs): records of 11 floats (dt ns, x1 y1 z1 x2 y2 z2 mm, q1 q2,
 * E1 E2 keV) -> cuts, energy smearing, element adjustment, scattergram,
 * LORs.  The smearing draws restate the library's documented stream:
 * Philox4x32-10 (Salmon et al., SC'11) under key (seed lo, seed hi), counter
 * (record lo, record hi, block, tag); uniforms from the top 24 bits; energy
 * block 0 tag 0xe6e7 (Box-Muller pairs (x, y) for E1, (z, w) for E2);
 * position blocks 1 and 2 tag 0x9051 (end 1: z = a.x, phi = a.y, r = b.x;
 * end 2: z = a.z, phi = a.w, r = b.y).  Same field layout as the library's
 * parameter structs. */
typedef struct {
  uint64_t events_min, events_max;
  int32_t has_events_min, has_events_max;
  int32_t has_charge_min, has_charge_max;
  float charge_min, charge_max;
  int32_t has_energy_min, has_energy_max;
  float energy_min, energy_max;
  float smear_energy_fwhm;
  int32_t has_z_max;
  float z_max_mm;
  float r_min_mm, dr_mm, dz_mm, da_mm;
  int32_t adjust;      /* 0 none, 1 centre, 2 random, 3 random z and phi */
  uint64_t seed;
} orc_ingest_t;

typedef struct {
  uint32_t phi_bins;
  uint32_t z_bins;  float z_length_mm;
  uint32_t dz_bins; float dz_max_mm;
  uint32_t r_bins;  float r_max_mm;
  uint32_t dt_bins; float dt_max_ps;
} orc_sgram_spec_t;

/* Returns the number of LORs written to out (capacity >= records in range);
 * *n_read = records in the event range. */
int64_t orc_ingest_f32(const float* records, size_t n_records, const orc_ingest_t* p,
                       const orc_sgram_spec_t* sgram_or_null, orc_lor_t* out, uint64_t* n_read);

/* LORs of a discretized detector (bin/make_sensitivity_image/discrete.rs:
 * 23-62, src/discrete.rs:150-171): element centres azimuth-major, pairs
 * (i, k) for i in 0..n, k in i..n, LOR from element k to element i, each end
 * adjusted (0 none, 1 centre, 2 random, 3 random z and phi) with draws from
 * Philox key (seed lo, seed hi), counter (pair lo, pair hi, block, 0x9a1b),
 * blocks 1 and 2 (end k: z = a.x, phi = a.y, r = b.x; end i: a.z, a.w,
 * b.y).  Writes min(pairs, capacity) LORs; returns the pair count and the
 * element count in *n_elements. */
int64_t orc_discrete_lors(float r_min_mm, float dr_mm, float dz_mm, float da_mm, float length_mm, int adjust,
                          uint64_t seed, orc_lor_t* out, size_t capacity, uint64_t* n_elements);

/* ---- Figures of merit (fom_oracle.cpp; src/fom.rs, src/bin/foms.rs) ----
 * Region kinds: 0 sphere (c = centre), 1 cylinder x (c = y, z), 2 cylinder
 * y (c = x, z), 3 cylinder z (c = x, y), 4 disc z (c = x, y, z).  Statistics
 * in binary32 in the reference's order (f64 = 0) or in binary64 (f64 = 1).
 * Return -1 when a region holds no voxel centre (the reference panics). */
typedef struct {
  int32_t kind;
  float c[3];
  float r_mm;
} orc_roi_t;
int orc_roi_stats(const float full_mm[3], const uint32_t n[3], const float* image, const orc_roi_t* rois,
                  size_t n_rois, int f64, uint64_t* counts, double* means, double* sigmas);
int orc_image_foms(const float full_mm[3], const uint32_t n[3], const float* image, const orc_roi_t* rois,
                   const float* activities, size_t n_rois, const orc_roi_t* bg_rois, size_t n_bg,
                   float bg_activity, int f64, float* crcs, float* snrs, float* bg_measured);
/* phantom 0 NEMA7, 1 Jaszczak; foms = 6 x (r, crc, bg variability, snr). */
int orc_phantom_foms(const float full_mm[3], const uint32_t n[3], const float* image, int phantom, int f64,
                     float foms[24], float* aocs, size_t aoc_capacity, size_t* n_aocs);

/* Philox4x32-10 block (for pinning the restatement against the paper's
 * known-answer vectors). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
