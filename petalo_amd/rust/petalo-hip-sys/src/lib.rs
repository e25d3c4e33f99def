//! Raw bindings to `libpetalo_hip.so`, the MI355X (gfx950) implementation of
//! petalo's reconstruction hot path: one `#[repr(C)]` struct per struct of
//! `include/petalo_hip.h`, field for field in the header's order, and one
//! `extern "C"` declaration per function.  `tests/test_rust_binding.py` checks
//! every struct's field names, offsets and size, every constant and every
//! function's name and parameter count against the header (compiled by the C
//! compiler), so that a Rust caller passes exactly the layout the library
//! reads.
//!
//! This follows the reference's own cdylib/FFI pattern (`mlem/src/lib.rs:1-6`).
//! A `petalo` build keeps its public API (`mlem::mlem`, `src/mlem.rs:8-31`;
//! `continuous::sensitivity_image`, `continuous.rs:3-17`;
//! `Projector::update_system_matrix_row`, `src/projectors/mod.rs:20-25`) and
//! calls these functions behind it.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

// ---- status codes ---------------------------------------------------------
pub const PETALO_OK: c_int = 0;
pub const PETALO_ERR_ARG: c_int = 1;
pub const PETALO_ERR_HIP: c_int = 2;
pub const PETALO_ERR_NOMEM: c_int = 3;
pub const PETALO_ERR_IO: c_int = 4;
pub const PETALO_ERR_PARSE: c_int = 5;
pub const PETALO_ERR_SIZE: c_int = 6;

// ---- options and enumerations ---------------------------------------------
pub const PETALO_SCHEME_AUTO: i32 = 0;
pub const PETALO_SCHEME_ATOMIC: i32 = 1;
pub const PETALO_SCHEME_TILE: i32 = 2;

pub const PETALO_PRECISION_F32: i32 = 0;
pub const PETALO_PRECISION_F64: i32 = 1;
pub const PETALO_PRECISION_F32_WALK: i32 = 2;

pub const PETALO_TIME_FORWARD_TILES: usize = 0;
pub const PETALO_TIME_FORWARD_REDUCE: usize = 1;
pub const PETALO_TIME_SCATTER: usize = 2;
pub const PETALO_TIME_UPDATE: usize = 3;
pub const PETALO_N_TIMES: usize = 4;

pub const PETALO_ADJUST_NONE: i32 = 0;
pub const PETALO_ADJUST_CENTRE: i32 = 1;
pub const PETALO_ADJUST_RANDOM: i32 = 2;
pub const PETALO_ADJUST_RANDOM_Z_PHI: i32 = 3;

pub const PETALO_ROI_SPHERE: i32 = 0;
pub const PETALO_ROI_CYLINDER_X: i32 = 1;
pub const PETALO_ROI_CYLINDER_Y: i32 = 2;
pub const PETALO_ROI_CYLINDER_Z: i32 = 3;
pub const PETALO_ROI_DISC_Z: i32 = 4;

pub const PETALO_PHANTOM_NEMA7: i32 = 0;
pub const PETALO_PHANTOM_JASZCZAK: i32 = 1;

// ---- structs ----------------------------------------------------------------
/// FOV::new's inputs (src/fov.rs:13-39).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_fov_t {
    pub full_size_mm: [f32; 3],
    pub n: [u32; 3],
}

/// The 32-byte LOR record (src/lor.rs:11-25): mm, ps, ratio.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_lor_t {
    pub p1: [f32; 3],
    pub p2: [f32; 3],
    pub dt_ps: f32,
    pub additive_correction: f32,
}

/// The [tof] table (src/config/mlem.rs:177-188).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_tof_t {
    pub enabled: i32,
    pub sigma_ps: f32,
    pub cutoff: f32,
}

/// Osem (src/mlem.rs:70-96), 1-based.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_osem_t {
    pub n_subsets: u64,
    pub iteration: u64,
    pub subset: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_stats_t {
    pub n_lors: u64,
    pub n_steps: u64,
    pub n_nz: u64,
    pub n_skipped: u64,
    pub n_missed: u64,
    pub n_runaway: u64,
    pub n_shifted: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_mlem_options_t {
    pub scheme: i32,
    pub sort_lors: i32,
    pub device: i32,
    pub tile: [i32; 3],
    pub autotune: i32,
    pub precision: i32,
    pub item_steps: i32,
}

/// Opaque MLEM/OSEM context.
#[repr(C)]
pub struct petalo_mlem_t {
    _private: [u8; 0],
}

/// Opaque batch of reconstructions.
#[repr(C)]
pub struct petalo_batch_t {
    _private: [u8; 0],
}

/// read_lors parameters (src/io/hdf5.rs:97-130).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_ingest_t {
    pub events_min: u64,
    pub events_max: u64,
    pub has_events_min: i32,
    pub has_events_max: i32,
    pub has_charge_min: i32,
    pub has_charge_max: i32,
    pub charge_min: f32,
    pub charge_max: f32,
    pub has_energy_min: i32,
    pub has_energy_max: i32,
    pub energy_min: f32,
    pub energy_max: f32,
    pub smear_energy_fwhm: f32,
    pub has_z_max: i32,
    pub z_max_mm: f32,
    pub r_min_mm: f32,
    pub dr_mm: f32,
    pub dz_mm: f32,
    pub da_mm: f32,
    pub adjust: i32,
    pub seed: u64,
}

/// [scatter_correction] axes (src/lorogram/build_scattergram.rs:60-108).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_scattergram_spec_t {
    pub phi_bins: u32,
    pub z_bins: u32,
    pub z_length_mm: f32,
    pub dz_bins: u32,
    pub dz_max_mm: f32,
    pub r_bins: u32,
    pub r_max_mm: f32,
    pub dt_bins: u32,
    pub dt_max_ps: f32,
}

/// The MLEM config (src/config/mlem.rs:73-253).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct petalo_config_t {
    pub input_file: [c_char; 512],
    pub input_dataset: [c_char; 256],
    pub iterations_number: u64,
    pub iterations_subsets: u64,
    pub fov: petalo_fov_t,
    pub has_tof: i32,
    pub tof: petalo_tof_t,
    pub has_attenuation_correction: i32,
    pub sensitivity_image: [c_char; 512],
    pub has_energy_min: i32,
    pub has_energy_max: i32,
    pub has_charge_min: i32,
    pub has_charge_max: i32,
    pub has_events_min: i32,
    pub has_events_max: i32,
    pub energy_min: f32,
    pub energy_max: f32,
    pub charge_min: f32,
    pub charge_max: f32,
    pub events_min: u64,
    pub events_max: u64,
    pub has_scatter_correction: i32,
    pub scatter: petalo_scattergram_spec_t,
    pub has_smear_energy: i32,
    pub smear_energy_fwhm: f32,
    pub has_detector_full_axial_length: i32,
    pub detector_full_axial_length_mm: f32,
}

/// A region of interest (src/fom.rs ROI).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_roi_t {
    pub kind: i32,
    pub c: [f32; 3],
    pub r_mm: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_roi_stats_t {
    pub count: u64,
    pub mean: f64,
    pub sigma: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct petalo_fom_t {
    pub r_mm: f32,
    pub crc: f32,
    pub bg_variability: f32,
    pub snr: f32,
}

// ---- functions ---------------------------------------------------------------
#[link(name = "petalo_hip")]
extern "C" {
    pub fn petalo_last_error() -> *const c_char;
    pub fn petalo_version() -> c_int;
    pub fn petalo_device_count(count: *mut c_int) -> c_int;
    pub fn petalo_device_synchronize(device: c_int) -> c_int;
    pub fn petalo_copy_probe(bytes: usize, iters: c_int, ms_per_copy: *mut f64) -> c_int;

    pub fn petalo_fov_derive(fov: *const petalo_fov_t, half_width: *mut f32, voxel_size: *mut f32) -> c_int;
    pub fn petalo_fov_parse(text: *const c_char, out: *mut petalo_fov_t) -> c_int;
    pub fn petalo_fov_voxel_centre(fov: *const petalo_fov_t, index3: *const u32, centre: *mut f32) -> c_int;

    pub fn petalo_mlem_create(
        out: *mut *mut petalo_mlem_t,
        fov: *const petalo_fov_t,
        tof_or_null: *const petalo_tof_t,
        lors: *const petalo_lor_t,
        n_lors: usize,
        sensitivity_or_null: *const f32,
        n_subsets: usize,
        options: *const petalo_mlem_options_t,
    ) -> c_int;
    pub fn petalo_mlem_next(ctx: *mut petalo_mlem_t, image_out_or_null: *mut f32, id_out: *mut petalo_osem_t) -> c_int;
    pub fn petalo_mlem_get_image(ctx: *mut petalo_mlem_t, image_out: *mut f32) -> c_int;
    pub fn petalo_mlem_get_image_f64(ctx: *mut petalo_mlem_t, image_out: *mut f64) -> c_int;
    pub fn petalo_mlem_set_image(ctx: *mut petalo_mlem_t, image: *const f32) -> c_int;
    pub fn petalo_mlem_walk_stats(ctx: *mut petalo_mlem_t, subset: usize, out: *mut petalo_stats_t) -> c_int;
    pub fn petalo_mlem_last_skipped(ctx: *mut petalo_mlem_t, skipped: *mut u64) -> c_int;
    pub fn petalo_mlem_accumulator(
        ctx: *mut petalo_mlem_t,
        unit_out: *mut f64,
        r_max_out: *mut f64,
        r_geomean_out: *mut f64,
    ) -> c_int;
    pub fn petalo_mlem_set_timing(ctx: *mut petalo_mlem_t, on: c_int) -> c_int;
    pub fn petalo_mlem_set_graphs(ctx: *mut petalo_mlem_t, on: c_int) -> c_int;
    pub fn petalo_mlem_last_timing(ctx: *mut petalo_mlem_t, project_ms: *mut f64, update_ms: *mut f64) -> c_int;
    pub fn petalo_mlem_last_phase_timing(
        ctx: *mut petalo_mlem_t,
        forward_ms: *mut f64,
        scatter_ms: *mut f64,
        update_ms: *mut f64,
    ) -> c_int;
    pub fn petalo_mlem_last_kernel_timing(ctx: *mut petalo_mlem_t, ms: *mut f64, n: usize) -> c_int;
    pub fn petalo_mlem_drain_kernel_timings(
        ctx: *mut petalo_mlem_t,
        ms: *mut f64,
        max_updates: usize,
        n_updates: *mut usize,
    ) -> c_int;
    pub fn petalo_mlem_info(ctx: *mut petalo_mlem_t, scheme: *mut i32, n_runs_total: *mut u64) -> c_int;
    pub fn petalo_mlem_tile_shape(ctx: *mut petalo_mlem_t, tile_out: *mut i32) -> c_int;
    pub fn petalo_mlem_work_items(
        ctx: *mut petalo_mlem_t,
        item_steps_out: *mut u64,
        n_items_total_out: *mut u64,
    ) -> c_int;
    pub fn petalo_mlem_destroy(ctx: *mut petalo_mlem_t) -> c_int;

    pub fn petalo_batch_create(
        out: *mut *mut petalo_batch_t,
        n: usize,
        fovs: *const petalo_fov_t,
        tofs_or_null: *const petalo_tof_t,
        lors: *const *const petalo_lor_t,
        n_lors: *const usize,
        sensitivities_or_null: *const *const f32,
        n_subsets: usize,
        options: *const petalo_mlem_options_t,
    ) -> c_int;
    pub fn petalo_batch_next(batch: *mut petalo_batch_t, id_out_or_null: *mut petalo_osem_t) -> c_int;
    pub fn petalo_batch_work_items(
        batch: *mut petalo_batch_t,
        item_steps_out: *mut u64,
        n_items_total_out: *mut u64,
    ) -> c_int;
    pub fn petalo_batch_get_image(batch: *mut petalo_batch_t, r: usize, image_out: *mut f32) -> c_int;
    pub fn petalo_batch_set_options(batch: *mut petalo_batch_t, timing: c_int, use_graphs: c_int) -> c_int;
    pub fn petalo_batch_set_schedule(batch: *mut petalo_batch_t, persistent: c_int) -> c_int;
    pub fn petalo_batch_last_ms(batch: *mut petalo_batch_t, ms: *mut f64) -> c_int;
    pub fn petalo_batch_destroy(batch: *mut petalo_batch_t) -> c_int;

    pub fn petalo_ingest_lors(
        records: *const f32,
        n_records: usize,
        params: *const petalo_ingest_t,
        scatter_or_null: *const petalo_scattergram_spec_t,
        out: *mut petalo_lor_t,
        capacity: usize,
        n_out: *mut usize,
        n_read_out: *mut u64,
    ) -> c_int;

    pub fn petalo_make_sensitivity_image(
        att_fov: *const petalo_fov_t,
        mu: *const f32,
        out_fov_or_null: *const petalo_fov_t,
        lors_or_null: *const petalo_lor_t,
        n_lors: usize,
        seed: u64,
        detector_length_mm: f32,
        detector_diameter_mm: f32,
        out: *mut f32,
        stats_or_null: *mut petalo_stats_t,
    ) -> c_int;
    pub fn petalo_make_sensitivity_image_discrete(
        att_fov: *const petalo_fov_t,
        mu: *const f32,
        out_fov_or_null: *const petalo_fov_t,
        detector_length_mm: f32,
        r_min_mm: f32,
        dr_mm: f32,
        dz_mm: f32,
        da_mm: f32,
        adjust: i32,
        seed: u64,
        out: *mut f32,
        stats_or_null: *mut petalo_stats_t,
    ) -> c_int;
    pub fn petalo_release_workspaces() -> c_int;
    pub fn petalo_sensitivity_image_noalloc(
        fov: *const petalo_fov_t,
        mu: *const f32,
        lors: *const petalo_lor_t,
        n_lors: usize,
        out: *mut f32,
        stream: *mut c_void,
    ) -> c_int;
    pub fn petalo_sensitivity_lors(
        att_fov: *const petalo_fov_t,
        n_lors: usize,
        seed: u64,
        detector_length_mm: f32,
        detector_diameter_mm: f32,
        out: *mut petalo_lor_t,
    ) -> c_int;
    pub fn petalo_density_to_attenuation(image: *mut f32, n_voxels: usize, rho_to_mu_cm2_per_g: f32) -> c_int;

    pub fn petalo_system_matrix_rows(
        fov: *const petalo_fov_t,
        tof_or_null: *const petalo_tof_t,
        lors: *const petalo_lor_t,
        n_lors: usize,
        cap: usize,
        idx: *mut i32,
        w: *mut f32,
        count: *mut u32,
        steps: *mut u32,
    ) -> c_int;

    pub fn petalo_image3d_write(path: *const c_char, fov: *const petalo_fov_t, data: *const f32) -> c_int;
    pub fn petalo_image3d_read_header(path: *const c_char, fov_out: *mut petalo_fov_t) -> c_int;
    pub fn petalo_image3d_read(path: *const c_char, fov_out: *mut petalo_fov_t, data: *mut f32, capacity: usize)
        -> c_int;

    pub fn petalo_config_parse(toml_text: *const c_char, out: *mut petalo_config_t) -> c_int;
    pub fn petalo_config_read(path: *const c_char, out: *mut petalo_config_t) -> c_int;
    pub fn petalo_config_ingest(
        config: *const petalo_config_t,
        r_min_mm: f32,
        dr_mm: f32,
        dz_mm: f32,
        da_mm: f32,
        seed: u64,
        out: *mut petalo_ingest_t,
    ) -> c_int;
    pub fn petalo_parse_quantity(text: *const c_char, kind: c_char, value: *mut f32) -> c_int;

    pub fn petalo_roi_stats(
        fov: *const petalo_fov_t,
        image: *const f32,
        rois: *const petalo_roi_t,
        n_rois: usize,
        out: *mut petalo_roi_stats_t,
    ) -> c_int;
    pub fn petalo_image_foms(
        fov: *const petalo_fov_t,
        image: *const f32,
        rois: *const petalo_roi_t,
        activities: *const f32,
        n_rois: usize,
        background_rois: *const petalo_roi_t,
        n_background: usize,
        background_activity: f32,
        crcs_out: *mut f32,
        snrs_out: *mut f32,
        background_measured_out: *mut f32,
    ) -> c_int;
    pub fn petalo_phantom_foms(
        fov: *const petalo_fov_t,
        image: *const f32,
        phantom: i32,
        foms_out: *mut petalo_fom_t,
        aocs_out: *mut f32,
        aoc_capacity: usize,
        n_aocs_out: *mut usize,
    ) -> c_int;

    pub fn petalo_synth_phantom_lors(seed: u64, n_lors: usize, tof_fwhm_ps: f32, out: *mut petalo_lor_t) -> c_int;
    pub fn petalo_philox4x32_10(ctr: *const u32, key: *const u32, out: *mut u32) -> c_int;
    pub fn petalo_synth_density(fov: *const petalo_fov_t, out: *mut f32) -> c_int;
}
