/*
 * petalo_hip.h -- C ABI of the MI355X (gfx950) implementation of petalo's
 * image-reconstruction hot path: the Siddon projector, its TOF-weighted
 * adjoint, the MLEM/OSEM multiplicative update and the sensitivity image.
 *
 * Conventions
 *  - Every function returns 0 (PETALO_OK) or a PETALO_ERR_* code; the
 *    message of the last failure on the calling thread is available from
 *    petalo_last_error().  Nothing unwinds across this boundary.  Contract
 *    violations that panic in the reference (wrong image length, mismatched
 *    sensitivity image) are reported as PETALO_ERR_SIZE.
 *  - Units are mm and ps (the reference's base units, units/src/lib.rs:36-47).
 *  - Host pointers in, host pointers out.  The library copies inputs to
 *    device memory; the caller keeps ownership of its arrays.
 *  - A context is single-threaded (one HIP stream each); separate contexts
 *    may be used from separate threads.  Calls that return host data
 *    synchronise before returning.
 *  - Image layout: x fastest, i = ix + (iy + iz*ny)*nx (src/index.rs:15-20).
 */
#ifndef PETALO_HIP_H
#define PETALO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PETALO_OK 0
#define PETALO_ERR_ARG 1     /* invalid argument                       */
#define PETALO_ERR_HIP 2     /* HIP runtime failure / no GPU            */
#define PETALO_ERR_NOMEM 3   /* allocation failed                       */
#define PETALO_ERR_IO 4      /* file read/write failed                  */
#define PETALO_ERR_PARSE 5   /* config / FOV string could not be parsed */
#define PETALO_ERR_SIZE 6    /* image size mismatch (reference panics)  */

/* Field of view: full extent and voxel counts.  The library derives
 * half_width = full/2 and voxel_size = (half_width*2)/n in binary32 exactly
 * as FOV::new does.  Replaces petalo::FOV (src/fov.rs:13-39). */
typedef struct {
  float full_size_mm[3];
  uint32_t n[3];
} petalo_fov_t;

/* Line of response, 32 bytes: endpoints (mm), dt = t2 - t1 (ps), and the
 * multiplicative additive-correction factor.  Replaces petalo::LOR
 * (src/lor.rs:11-25). */
typedef struct {
  float p1[3];
  float p2[3];
  float dt_ps;
  float additive_correction;
} petalo_lor_t;

/* TOF parameters as in the [tof] config table (src/config/mlem.rs:177-188):
 * sigma is the time sigma in ps (the kernel width is sigma*C), cutoff in
 * sigmas (default 3).  Replaces Option<Tof> / Siddon::new (siddon.rs:47). */
typedef struct {
  int32_t enabled;
  float sigma_ps;
  float cutoff;
} petalo_tof_t;

/* OSEM progress counter, 1-based (src/mlem.rs:70-96). */
typedef struct {
  uint64_t n_subsets;
  uint64_t iteration;
  uint64_t subset;
} petalo_osem_t;

/* Work counters of the last projection pass. */
typedef struct {
  uint64_t n_lors;     /* LORs processed                                     */
  uint64_t n_steps;    /* iterations of the reference walk loop (full chord) */
  uint64_t n_nz;       /* pushed system-matrix elements (weight > 0)         */
  uint64_t n_skipped;  /* LORs dropped by the index check (projector.rs:103-118) */
  uint64_t n_missed;   /* LORs missing the FOV                               */
  uint64_t n_runaway;  /* walks stopped at the step cap                      */
  uint64_t n_shifted;  /* walks that start one voxel short of their FOV entry:
                          an entry coordinate in voxel units still negative
                          after the reference's 1e-5 snap (siddon.rs:134-153),
                          which floor + the saturating cast (:312-313) put in
                          voxel 0; the reference's own behaviour, counted     */
} petalo_stats_t;

/* Backprojection schemes for petalo_mlem_options_t.scheme. */
#define PETALO_SCHEME_AUTO 0     /* tiled, or autotuned (options.autotune)  */
#define PETALO_SCHEME_ATOMIC 1   /* plain: one lane per LOR, f32 global atomics */
#define PETALO_SCHEME_TILE 2     /* image tiles in LDS, binary64 accumulation */

/* Arithmetic of a context (petalo_mlem_options_t.precision). */
#define PETALO_PRECISION_F32 0       /* the reference's binary32 walk and arithmetic */
#define PETALO_PRECISION_F64 1       /* everything in binary64: walk setup, walk,
                                        TOF weights, projections, image, update   */
#define PETALO_PRECISION_F32_WALK 2  /* the binary32 walk's discrete decisions
                                        (first voxel, start fraction, crossed axis
                                        of every step, pushes, end) with binary64
                                        weights, projections, image and update    */

typedef struct {
  int32_t scheme;        /* PETALO_SCHEME_*                                   */
  int32_t sort_lors;     /* 1: reorder LORs inside each subset for locality   */
  int32_t device;        /* HIP device ordinal                                */
  int32_t tile[3];       /* tiled scheme: tile shape in voxels (x, y, z); a
                            zero takes the default (x, y: about 40, the axis
                            split evenly; z: as many slices as fit 10240 LDS
                            entries).  At most 20416 LDS entries per tile,
                            rows padded to an odd pitch.                      */
  int32_t autotune;      /* with scheme AUTO: 1 times the default tile, the
                            atomic scheme and tiles of about 24, 32, 48 and
                            64 voxels across on up to 2^21 LORs of the first
                            subset, and keeps the default unless another is
                            at least 3% faster (creation takes about a second
                            longer), then, with item_steps 0, the work
                            granularity (below); 0 takes the tiled scheme
                            with the tile given or defaulted.                */
  int32_t precision;     /* PETALO_PRECISION_*; the binary64 modes run one lane
                            per LOR with binary64 global atomics whatever the
                            scheme (reference modes for precision studies;
                            not with autotune, not in batches)               */
  int32_t item_steps;    /* tiled scheme: replayed steps per work item (one
                            workgroup each; at least 8192); 0 takes the
                            automatic rule (about 2048 items per subset
                            update, at least 16 steps per LDS entry of the
                            tile), and with autotune the tuner then also
                            times 2^14 .. 2^20 on the context's first
                            subsets and keeps the rule unless one is 3%
                            faster                                           */
} petalo_mlem_options_t;

/* ---- library ----------------------------------------------------------- */
const char* petalo_last_error(void);
int petalo_version(void);
/* Number of visible HIP devices (0 without a GPU; never fails for that). */
int petalo_device_count(int* count);
/* Blocks until all work on the given device has finished. */
int petalo_device_synchronize(int device);
/* Calibration probe for hardware counters: copies `bytes` between two device
 * buffers with a plain streaming kernel (k_copy_probe) iters+1 times and
 * returns the average time of the last iters copies. */
int petalo_copy_probe(size_t bytes, int iters, double* ms_per_copy);

/* ---- FOV (src/fov.rs:22-48,75-91) ------------------------------------- */
int petalo_fov_derive(const petalo_fov_t* fov, float half_width[3], float voxel_size[3]);
/* Parses '1 mm  2 mm  0.2 cm  100 100 100' (FOV::from_str, src/fov.rs:75-91). */
int petalo_fov_parse(const char* text, petalo_fov_t* out);
/* Centre of voxel (ix,iy,iz) (FOV::voxel_centre, src/fov.rs:42-48). */
int petalo_fov_voxel_centre(const petalo_fov_t* fov, const uint32_t index3[3], float centre[3]);

/* ---- MLEM / OSEM iterator (replaces petalo::mlem::mlem, src/mlem.rs:8-31) */
typedef struct petalo_mlem petalo_mlem_t;

/* Copies LORs (and the sensitivity image, NULL = ones) to the device and
 * initialises the image to ones (mlem.rs:17-19).  options may be NULL. */
int petalo_mlem_create(petalo_mlem_t** out, const petalo_fov_t* fov, const petalo_tof_t* tof_or_null,
                       const petalo_lor_t* lors, size_t n_lors, const float* sensitivity_or_null,
                       size_t n_subsets, const petalo_mlem_options_t* options);
/* One subset update = one Iterator::next (mlem.rs:25-30).  id_out receives
 * the Osem counter *before* it advances; image_out (N_vox floats, may be
 * NULL) receives the updated image. */
int petalo_mlem_next(petalo_mlem_t* ctx, float* image_out_or_null, petalo_osem_t* id_out);
/* Copies the current image to the host (a binary64 image rounded). */
int petalo_mlem_get_image(petalo_mlem_t* ctx, float* image_out);
/* The current image in binary64 (binary64 precision modes only). */
int petalo_mlem_get_image_f64(petalo_mlem_t* ctx, double* image_out);
/* Replaces the current image (for resuming from a written subset image). */
int petalo_mlem_set_image(petalo_mlem_t* ctx, const float* image);
/* Exact walk statistics of the LOR set (full reference walk per LOR), for
 * subset `subset` (1-based; 0 = all LORs). */
int petalo_mlem_walk_stats(petalo_mlem_t* ctx, size_t subset, petalo_stats_t* out);
/* LORs skipped by the index check during the last update. */
int petalo_mlem_last_skipped(petalo_mlem_t* ctx, uint64_t* skipped);
/* The last tiled update's backprojection accumulator: the value of one unit
 * of its first work item's 64-bit fixed-point tile sums, or 0 where it summed
 * in binary64 (some r = 1/(p*ac) negative or not finite, or the largest r too
 * far above the geometric mean of the positive r for the item's unit), with
 * the update's largest r and that geometric mean (either pointer may be
 * NULL; NaN where no LOR had a positive r).  No counterpart in the reference
 * (which sums in binary32); for tests and diagnostics. */
int petalo_mlem_accumulator(petalo_mlem_t* ctx, double* unit_out, double* r_max_out, double* r_geomean_out);
/* Per-kernel device time of the last update in ms, measured with HIP events
 * on the context's stream when timing is on. */
int petalo_mlem_set_timing(petalo_mlem_t* ctx, int on);
/* Without timing, each subset update is captured once as a HIP graph and
 * replayed (default on); off launches the kernels one by one. */
int petalo_mlem_set_graphs(petalo_mlem_t* ctx, int on);
int petalo_mlem_last_timing(petalo_mlem_t* ctx, double* project_ms, double* update_ms);
/* Phase split of the last update: forward projection, backprojection
 * scatter (0 for the atomic scheme, which does both in one kernel) and the
 * image update. */
int petalo_mlem_last_phase_timing(petalo_mlem_t* ctx, double* forward_ms, double* scatter_ms, double* update_ms);
/* Kernel-level split of the last update, written to ms[0 .. min(n, 4)):
 *   FORWARD_TILES   tiled scheme: the per-tile forward kernel;
 *                   atomic scheme: the fused projection kernel
 *   FORWARD_REDUCE  per-LOR reduction to 1/(p*ac) plus the one-lane path
 *                   for unbounded walks (tiled scheme; 0 otherwise)
 *   SCATTER         tiled backprojection (0 for the atomic scheme)
 *   UPDATE          image update */
enum {
  PETALO_TIME_FORWARD_TILES = 0,
  PETALO_TIME_FORWARD_REDUCE = 1,
  PETALO_TIME_SCATTER = 2,
  PETALO_TIME_UPDATE = 3,
  PETALO_N_TIMES = 4
};
int petalo_mlem_last_kernel_timing(petalo_mlem_t* ctx, double* ms, size_t n);
/* The same PETALO_N_TIMES timings for each of the most recent timed subset
 * updates since timing was switched on or last drained (at most
 * max_updates, and at most the 256 the context keeps), oldest first, into
 * ms[u * PETALO_N_TIMES + k]; *n_updates is how many.  Waits for the
 * context's stream.  Timed updates do not wait for each other, so a loop of
 * them runs back to back on the device.  No counterpart in the reference. */
int petalo_mlem_drain_kernel_timings(petalo_mlem_t* ctx, double* ms, size_t max_updates, size_t* n_updates);
/* Scheme in use and, for the tiled scheme, the number of tile runs. */
int petalo_mlem_info(petalo_mlem_t* ctx, int32_t* scheme, uint64_t* n_runs_total);
/* The tile shape of a tiled context (zeros for the atomic scheme): after
 * autotuning, the shape it chose. */
int petalo_mlem_tile_shape(petalo_mlem_t* ctx, int32_t tile_out[3]);
/* The work granularity of a tiled context: its replayed steps per work item
 * (0: the automatic rule; after autotuning, the choice) and its work items
 * over all subsets. */
int petalo_mlem_work_items(petalo_mlem_t* ctx, uint64_t* item_steps_out, uint64_t* n_items_total_out);
int petalo_mlem_destroy(petalo_mlem_t* ctx);

/* ---- Batched reconstructions -------------------------------------------
 * n independent MLEM/OSEM reconstructions (each as petalo_mlem_create: its
 * own FOV, TOF, LORs, sensitivity image), all with n_subsets subsets, advanced
 * together: each petalo_batch_next performs one subset update of every
 * reconstruction with one forward and one backprojection launch for all of
 * them (work items of all their image tiles, largest first), replayed as a
 * HIP graph per subset.  tofs_or_null: n entries or NULL (no TOF);
 * sensitivities_or_null: n pointers (entries may be NULL) or NULL.  Tiled
 * scheme only; options.tile applies to every reconstruction, options.item_steps
 * to every reconstruction's work items and to their merging, and
 * options.autotune (with item_steps 0) times the work granularity as for one
 * context, on the whole batch.  No counterpart in the reference, whose mlem
 * runs one reconstruction per call (src/mlem.rs:8-31). */
typedef struct petalo_batch petalo_batch_t;
int petalo_batch_create(petalo_batch_t** out, size_t n, const petalo_fov_t* fovs, const petalo_tof_t* tofs_or_null,
                        const petalo_lor_t* const* lors, const size_t* n_lors,
                        const float* const* sensitivities_or_null, size_t n_subsets,
                        const petalo_mlem_options_t* options);
/* One subset update of every reconstruction; id_out: the Osem before advance. */
int petalo_batch_next(petalo_batch_t* batch, petalo_osem_t* id_out_or_null);
/* The batch's replayed steps per merged work item (0: the automatic rule;
 * after autotuning, the choice) and its work items over all subsets. */
int petalo_batch_work_items(petalo_batch_t* batch, uint64_t* item_steps_out, uint64_t* n_items_total_out);
int petalo_batch_get_image(petalo_batch_t* batch, size_t r, float* image_out);
/* timing: HIP events around each petalo_batch_next (petalo_batch_last_ms);
 * use_graphs: replay captured graphs (default 1) or launch directly. */
int petalo_batch_set_options(petalo_batch_t* batch, int timing, int use_graphs);
/* persistent = 1: each subset update of the batch is one persistent launch
 * whose workgroups take every reconstruction's forward, reduction,
 * backprojection and update tasks from an atomic queue (each waiting, time
 * bounded, for the previous phase of its own reconstruction); 0: one launch
 * per phase.  No counterpart in the reference (one mlem iterator per call,
 * src/mlem.rs:8-31).  A wait past its limit makes petalo_batch_get_image
 * fail with PETALO_ERR_HIP. */
int petalo_batch_set_schedule(petalo_batch_t* batch, int persistent);
int petalo_batch_last_ms(petalo_batch_t* batch, double* ms);
int petalo_batch_destroy(petalo_batch_t* batch);

/* ---- LOR ingestion (src/io/hdf5.rs:97-228, src/discrete.rs:60-213,
 * src/lorogram.rs, src/lorogram/build_scattergram.rs) ----------------------
 * Turns the records of a reco_info/lors dataset, as the caller read them from
 * the HDF5 file (Hdf5Lor, hdf5.rs:282-297: dt [ns], x1 y1 z1 x2 y2 z2 [mm],
 * q1 q2, E1 E2 [keV]; 11 floats per record), into MLEM LORs on the device:
 * records [events_min, events_max) (hdf5.rs:36-66), cut on |z| < z_max, on
 * charge and on energy before smearing (511 keV * (1 - 1.6 fwhm)), energies
 * smeared with a Gaussian of FWHM fwhm * E (hdf5.rs:230-238), cut on the energy
 * bounds, both ends moved inside their detector element (discrete.rs:78-118,
 * `adjust`), optionally counted into a scattergram of trues (min E >= 510 keV)
 * and scatters (hdf5.rs:69-92), then converted (dt in ps; additive correction
 * = (scatters + trues) / trues of the LOR's bin, or FLT_MAX for an empty bin,
 * lorogram.rs:65-74; 1 without a scattergram).  The smearing draws come from a
 * Philox4x32-10 stream keyed by `seed` and counted by record index (the
 * reference uses an unseeded thread RNG). */
#define PETALO_ADJUST_NONE 0          /* Adjust::No                      */
#define PETALO_ADJUST_CENTRE 1        /* Adjust::Centre                  */
#define PETALO_ADJUST_RANDOM 2        /* Adjust::Random                  */
#define PETALO_ADJUST_RANDOM_Z_PHI 3  /* Adjust::RandomZPhi (the reader's) */

typedef struct {
  uint64_t events_min, events_max;   /* [input] events: record range [min, max) */
  int32_t has_events_min, has_events_max;
  int32_t has_charge_min, has_charge_max;  /* [input] charge: min <= q < max */
  float charge_min, charge_max;
  int32_t has_energy_min, has_energy_max;  /* [input] energy (after smearing) */
  float energy_min, energy_max;
  float smear_energy_fwhm;           /* [smear_energy] fwhm, a fraction of E */
  int32_t has_z_max;                 /* [detector_full_axial_length]: z_max = dz / 2 */
  float z_max_mm;
  float r_min_mm, dr_mm, dz_mm, da_mm;  /* element discretization (the dataset's
                                           r_min, dr, dz, da attributes) */
  int32_t adjust;                    /* PETALO_ADJUST_*                   */
  uint64_t seed;
} petalo_ingest_t;

/* Scattergram axes (config [scatter_correction]): bins = 0 and max = 0 leave
 * an axis out (one bin spanning everything); bins alone take the builder's
 * default range (r 30 mm, z 200 mm, dz 1000 mm, dt 1700 ps), a range alone 30
 * bins (build_scattergram.rs:60-108).  All axes out: no scattergram. */
typedef struct {
  uint32_t phi_bins;
  uint32_t z_bins;  float z_length_mm;
  uint32_t dz_bins; float dz_max_mm;
  uint32_t r_bins;  float r_max_mm;
  uint32_t dt_bins; float dt_max_ps;
} petalo_scattergram_spec_t;

/* out must hold capacity LORs (at most the number of records in the event
 * range are produced); *n_out = LORs written, *n_read_out = records read. */
int petalo_ingest_lors(const float* records, size_t n_records, const petalo_ingest_t* params,
                       const petalo_scattergram_spec_t* scatter_or_null, petalo_lor_t* out, size_t capacity,
                       size_t* n_out, uint64_t* n_read_out);

/* ---- Sensitivity image (continuous.rs:3-47, main.rs:77) ---------------
 * Backprojects n_lors LORs through the attenuation image mu [1/mm] on
 * att_fov, adapting each projection p to exp(-p) (projector.rs:72-74), into
 * out_fov (NULL = att_fov), normalised by n_lors.  With lors == NULL the
 * LORs are generated on the device: both ends uniform on the detector
 * cylinder (length, diameter), rejected until they enter att_fov, from a
 * Philox4x32-10 stream keyed by seed and counted by LOR index (attempt k of
 * LOR i: block (i, k, 0, 0x5e45); z = L (u0 - 1/2), theta = TAU u1 for p1,
 * u2, u3 for p2; u = top 24 bits / 2^24).  A LOR that finds no FOV-crossing
 * pair within 4096 attempts fails the call with PETALO_ERR_ARG (the reference
 * draws forever).  Accumulation is in FP64.  A non-NULL stats also counts
 * every step and pushed step of every LOR, which costs one more walk over
 * all of them; with NULL the binning jumps from tile to tile. */
int petalo_make_sensitivity_image(const petalo_fov_t* att_fov, const float* mu,
                                  const petalo_fov_t* out_fov_or_null, const petalo_lor_t* lors_or_null,
                                  size_t n_lors, uint64_t seed, float detector_length_mm,
                                  float detector_diameter_mm, float* out, petalo_stats_t* stats_or_null);
/* Sensitivity image of a discretized detector (discrete::sensitivity_image,
 * src/bin/make_sensitivity_image/discrete.rs:3-62): element centres of a
 * ring of inner radius r_min, thickness dr, element width dz x da (at
 * r_min + dr/2), axial length detector_length (discrete.rs:150-171); one LOR
 * per element pair i <= k, self pairs included, each end adjusted inside its
 * element (PETALO_ADJUST_*, draws from Philox keyed by seed and counted by
 * pair index); projected as the continuous image; normalised by n (n - 1) / 2
 * over the n elements (discrete.rs:16-20, which counts neither self pairs nor
 * the FOV filter).  stats->n_lors = pairs; a non-NULL stats costs a walk over
 * every step, as above. */
int petalo_make_sensitivity_image_discrete(const petalo_fov_t* att_fov, const float* mu,
                                           const petalo_fov_t* out_fov_or_null, float detector_length_mm,
                                           float r_min_mm, float dr_mm, float dz_mm, float da_mm, int32_t adjust,
                                           uint64_t seed, float* out, petalo_stats_t* stats_or_null);
/* The two calls above keep their device buffers and stream (about 10 GB for
 * 2^25-LOR batches at 120^3: the LORs, their runs and the scans) per device between
 * calls, which saves about 4 ms of allocation and release per call.  No
 * counterpart in the reference (make_sensitivity_image runs once per
 * process there): frees every kept set, on every device.  Calls running at
 * the same time keep theirs. */
int petalo_release_workspaces(void);
/* petalo_make_sensitivity_image for caller-given LORs on one FOV, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * mu (N_vox floats), lors (n_lors records) and out (N_vox floats) are DEVICE
 * pointers on the current device; stream is a hipStream_t (NULL: the default
 * stream) on which the work is enqueued and not waited for.  The binary64
 * accumulator and the counters live in the library's static device scratch
 * (one per device, kept until the process ends): images of at most 2^21
 * voxels (PETALO_ERR_SIZE beyond).  Calls that overlap on one device race on
 * that scratch.  Each LOR is projected by the one-lane kernel (skip and miss
 * rules as petalo_make_sensitivity_image); no statistics are returned. */
int petalo_sensitivity_image_noalloc(const petalo_fov_t* fov, const float* mu, const petalo_lor_t* lors,
                                     size_t n_lors, float* out, void* stream);
/* The LORs petalo_make_sensitivity_image generates for (seed, fov, cylinder). */
int petalo_sensitivity_lors(const petalo_fov_t* att_fov, size_t n_lors, uint64_t seed,
                            float detector_length_mm, float detector_diameter_mm, petalo_lor_t* out);
/* density [kg/m^3] -> mu [1/mm] in place, rho_to_mu in cm^2/g
 * (density_image_into_attenuation_image, main.rs:11-17,80-94). */
int petalo_density_to_attenuation(float* image, size_t n_voxels, float rho_to_mu_cm2_per_g);

/* ---- System-matrix rows (Siddon::new_system_matrix_row, siddon.rs:69-73)
 * For each LOR i: up to cap elements at idx/w[i*cap ...], count[i] = number
 * of pushed elements (may exceed cap: then only cap are written), steps[i] =
 * walk-loop iterations (UINT32_MAX for a walk stopped at the step cap). */
int petalo_system_matrix_rows(const petalo_fov_t* fov, const petalo_tof_t* tof_or_null,
                              const petalo_lor_t* lors, size_t n_lors, size_t cap, int32_t* idx,
                              float* w, uint32_t* count, uint32_t* steps);

/* ---- Image3D raw format (src/io/raw.rs:78-126) -------------------------
 * Big-endian: u16 n[3], f32 full_size_mm[3], then f32 data. */
int petalo_image3d_write(const char* path, const petalo_fov_t* fov, const float* data);
int petalo_image3d_read_header(const char* path, petalo_fov_t* fov_out);
int petalo_image3d_read(const char* path, petalo_fov_t* fov_out, float* data, size_t capacity);

/* ---- MLEM config TOML (src/config/mlem.rs:73-253) ----------------------
 * Tables: [input] (file, dataset, energy/charge/events bounds),
 * [iterations], [fov], [tof], [attenuation_correction], [scatter_correction]
 * (axes as petalo_scattergram_spec_t; all zero when absent), [smear_energy]
 * (fwhm as a ratio) and [detector_full_axial_length] (dz in mm).  Unknown
 * keys are rejected (deny_unknown_fields).  petalo_config_ingest maps a
 * config onto the ingestion parameters as io::hdf5::read_lors does
 * (hdf5.rs:97-130, 240-262). */
typedef struct {
  char input_file[512];
  char input_dataset[256];
  uint64_t iterations_number;
  uint64_t iterations_subsets;
  petalo_fov_t fov;
  int32_t has_tof;
  petalo_tof_t tof;
  int32_t has_attenuation_correction;
  char sensitivity_image[512];
  int32_t has_energy_min, has_energy_max, has_charge_min, has_charge_max, has_events_min, has_events_max;
  float energy_min, energy_max, charge_min, charge_max;
  uint64_t events_min, events_max;
  int32_t has_scatter_correction;
  petalo_scattergram_spec_t scatter;
  int32_t has_smear_energy;
  float smear_energy_fwhm;
  int32_t has_detector_full_axial_length;
  float detector_full_axial_length_mm;
} petalo_config_t;

int petalo_config_parse(const char* toml_text, petalo_config_t* out);
int petalo_config_read(const char* path, petalo_config_t* out);
/* Ingestion parameters of read_lors for this config: event range, charge and
 * energy bounds, smearing FWHM (the reference requires [smear_energy]:
 * PETALO_ERR_ARG without it), |z| < dz/2 of [detector_full_axial_length],
 * element discretization (r_min, dr, dz, da — the dataset's attributes) with
 * the RandomZPhi adjustment (hdf5.rs:247-253), the smearing seed. */
int petalo_config_ingest(const petalo_config_t* config, float r_min_mm, float dr_mm, float dz_mm, float da_mm,
                         uint64_t seed, petalo_ingest_t* out);
/* Parses a quantity string such as "200 ps", "3 cm", "1.5 mm" into base
 * units (mm or ps); kind: 'L' length, 'T' time. */
int petalo_parse_quantity(const char* text, char kind, float* value);

/* ---- Figures of merit (src/fom.rs:20-461, src/bin/foms.rs) --------------
 * Regions of interest (fom.rs ROI): a voxel belongs to a region when its
 * centre (fov.rs:42-53) satisfies, in binary32 as written (fom.rs:34-63):
 *   SPHERE      (x-c0)^2 + (y-c1)^2 + (z-c2)^2 < r^2
 *   CYLINDER_X  (y-c0)^2 + (z-c1)^2 < r^2
 *   CYLINDER_Y  (x-c0)^2 + (z-c1)^2 < r^2
 *   CYLINDER_Z  (x-c0)^2 + (y-c1)^2 < r^2
 *   DISC_Z      z == c2 exactly and (x-c0)^2 + (y-c1)^2 < r^2 */
#define PETALO_ROI_SPHERE     0
#define PETALO_ROI_CYLINDER_X 1
#define PETALO_ROI_CYLINDER_Y 2
#define PETALO_ROI_CYLINDER_Z 3
#define PETALO_ROI_DISC_Z     4
typedef struct {
  int32_t kind;  /* PETALO_ROI_* */
  float c[3];    /* centre coordinates as listed above, mm (unused ones ignored) */
  float r_mm;
} petalo_roi_t;
/* Per region: voxel count, mean and population standard deviation of the
 * voxel values (fom::mean, fom::mu_and_sigma, fom.rs:276-288), accumulated
 * in FP64 (the reference sums binary32 values in voxel order). */
typedef struct {
  uint64_t count;
  double mean;
  double sigma;
} petalo_roi_stats_t;
/* Statistics of n regions of an image (host pointers); one workgroup per
 * region over the region's bounding box of voxels.  An empty region has
 * count 0 and mean = sigma = NaN. */
int petalo_roi_stats(const petalo_fov_t* fov, const float* image, const petalo_roi_t* rois, size_t n_rois,
                     petalo_roi_stats_t* out);
/* Image::foms (fom.rs:426-450): background = mean over background regions of
 * their means; per region crc (fom.rs:452-461, percent, hot when the
 * region's activity exceeds the background's) and snr = (mean - background) /
 * sigma.  PETALO_ERR_ARG if a region holds no voxel (the reference panics). */
int petalo_image_foms(const petalo_fov_t* fov, const float* image, const petalo_roi_t* rois,
                      const float* activities, size_t n_rois, const petalo_roi_t* background_rois,
                      size_t n_background, float background_activity, float* crcs_out, float* snrs_out,
                      float* background_measured_out);
/* The foms tool's phantom analyses (src/bin/foms.rs): six hot spheres (4:1)
 * in discs on the slice of the sphere centres, background discs of each
 * sphere's radius on the phantom's background slices; per sphere (in the
 * reference's order) radius, crc %, background variability % and snr %.
 * NEMA7 also gives the lung-insert accuracy of corrections per slice
 * (foms.rs 7.4.2: mean in a 15 mm disc / the 37 mm sphere's crc, x 100),
 * written up to aoc_capacity values, *n_aocs_out = how many there are. */
#define PETALO_PHANTOM_NEMA7    0
#define PETALO_PHANTOM_JASZCZAK 1
typedef struct {
  float r_mm;
  float crc;
  float bg_variability;
  float snr;
} petalo_fom_t;
int petalo_phantom_foms(const petalo_fov_t* fov, const float* image, int32_t phantom, petalo_fom_t foms_out[6],
                        float* aocs_out, size_t aoc_capacity, size_t* n_aocs_out);

/* ---- Synthetic workloads (seeded, Philox4x32-10 counter-based) ---------
 * Water cylinder (r 80 mm, |z| <= 75 mm, activity 1) with six hot spheres
 * (activity 4, radii 5..15 mm on a 50 mm ring), isotropic back-to-back
 * emission, both ends on the detector cylinder (r 350 mm, |z| <= 150 mm);
 * TOF dt with Gaussian timing noise of the given FWHM (0 = exact dt).
 * Deterministic for a given (seed, n) regardless of thread count. */
int petalo_synth_phantom_lors(uint64_t seed, size_t n_lors, float tof_fwhm_ps, petalo_lor_t* out);
/* The counter-based generator behind the synthetic and sensitivity LOR
 * streams: Philox4x32-10 (Salmon et al., SC'11) of counter ctr under key. */
int petalo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Density map [kg/m^3] for the attenuation workload: water cylinder
 * (1000) with a bone insert (1791.67) of r 15 mm at (-40, 0), |z| <= 75. */
int petalo_synth_density(const petalo_fov_t* fov, float* out);

#ifdef __cplusplus
}
#endif
#endif
