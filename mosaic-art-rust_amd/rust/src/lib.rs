//! mosaic-art-rust's tile-matching API on the gfx950 HIP path.
//!
//! The types and methods keep the names, argument meaning and results of the
//! reference library (paths relative to the mosaic-art-rust repository):
//!   * `SimilarityDatabase`           src/similarity.rs:30-137
//!   * `AdjacencyPenaltyCalculator`   src/adjacency.rs:49-214
//!   * `MosaicOptimizer`, `OptimizationConfig`, `OptimizationResult`
//!                                    src/optimizer.rs:7-221
//!   * `ColorAdjustment`, `calculate_optimal_adjustment`
//!                                    src/color_adjustment.rs:6-154
//!   * `calculate_average_lab`        src/lib.rs:94-119
//!   * `generate_mosaic`              src/main.rs:477-685 minus file I/O
//! Material identity stays a `PathBuf` on this side; the library sees row
//! indices in insertion order, as `SimilarityDatabase::path_to_index` keeps
//! them.  Grids are the reference's `[Vec<Option<PathBuf>>]`, row-major.
//! Images are RGB8 slices (`width * height * 3` bytes, row-major).
//!
//! A `Context` owns one HIP device and stream; it is not `Send` (one call in
//! flight per context), so the handles here share it through `Rc`.
pub mod hip_ffi;

use hip_ffi as ffi;
use std::collections::HashMap;
use std::path::{Path, PathBuf};
use std::ptr;
use std::rc::Rc;

/// An `MH_ERR_*` code with the library's message.
#[derive(Debug, Clone)]
pub struct Error {
    pub code: i32,
    pub message: String,
}

impl std::fmt::Display for Error {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "mosaic_hip error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for Error {}

pub type Result<T> = std::result::Result<T, Error>;

/// One HIP device and stream (`mh_context`).
pub struct Context {
    raw: *mut ffi::mh_context,
}

impl Context {
    pub fn new(device: i32) -> Result<Rc<Context>> {
        let mut raw = ptr::null_mut();
        ffi::check(unsafe { ffi::mh_context_create(device, &mut raw) })?;
        Ok(Rc::new(Context { raw }))
    }

    pub fn raw(&self) -> *mut ffi::mh_context {
        self.raw
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { ffi::mh_context_destroy(self.raw) };
    }
}

/// CIE L*a*b* (D65), f32 as palette's `Lab`.
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct Lab {
    pub l: f32,
    pub a: f32,
    pub b: f32,
}

impl Lab {
    pub fn new(l: f32, a: f32, b: f32) -> Self {
        Lab { l, a, b }
    }
}

/// `MosaicGeneratorImpl::calculate_average_lab` (src/lib.rs:94-119): the f32
/// mean of every pixel's Lab, summed in row-major order.
pub fn calculate_average_lab(ctx: &Context, rgb: &[u8], width: u32, height: u32) -> Result<Lab> {
    assert_eq!(rgb.len(), (width * height * 3) as usize);
    let mut out = [0f32; 3];
    ffi::check(unsafe {
        ffi::mh_average_lab(ctx.raw, rgb.as_ptr(), width as i32, height as i32, out.as_mut_ptr())
    })?;
    Ok(Lab::new(out[0], out[1], out[2]))
}

// ------------------------------------------------------- similarity database
/// `SimilarityDatabase` (src/similarity.rs:30-137).  The Lab table and the
/// similarity matrix live on the GPU once `build_similarities` has run.
pub struct SimilarityDatabase {
    ctx: Rc<Context>,
    path_to_index: HashMap<PathBuf, usize>,
    index_to_path: HashMap<usize, PathBuf>,
    lab_colors: Vec<Lab>,
    lib: *mut ffi::mh_library,
    built: bool,
}

impl SimilarityDatabase {
    pub fn new(ctx: Rc<Context>) -> Self {
        SimilarityDatabase {
            ctx,
            path_to_index: HashMap::new(),
            index_to_path: HashMap::new(),
            lab_colors: Vec::new(),
            lib: ptr::null_mut(),
            built: false,
        }
    }

    /// src/similarity.rs:61-66: the tile gets the next index.
    pub fn add_tile(&mut self, path: PathBuf, lab_color: Lab) {
        let index = self.lab_colors.len();
        self.path_to_index.insert(path.clone(), index);
        self.index_to_path.insert(index, path);
        self.lab_colors.push(lab_color);
        self.built = false;
    }

    fn upload(&mut self) -> Result<()> {
        if !self.lib.is_null() {
            unsafe { ffi::mh_library_destroy(self.lib) };
            self.lib = ptr::null_mut();
        }
        let flat: Vec<f32> = self.lab_colors.iter().flat_map(|c| [c.l, c.a, c.b]).collect();
        ffi::check(unsafe {
            ffi::mh_library_from_lab(self.ctx.raw, flat.as_ptr(), self.lab_colors.len() as i32,
                                     &mut self.lib)
        })
    }

    /// src/similarity.rs:69-82: the n x n distances, resident in HBM.
    pub fn build_similarities(&mut self) -> Result<()> {
        self.upload()?;
        ffi::check(unsafe { ffi::mh_library_build_similarities(self.lib) })?;
        self.built = true;
        Ok(())
    }

    /// src/similarity.rs:85-103: 0.0 on the diagonal, None for an unknown
    /// path or before build_similarities.
    pub fn get_similarity(&self, path1: &Path, path2: &Path) -> Option<f32> {
        let i = *self.path_to_index.get(path1)?;
        let j = *self.path_to_index.get(path2)?;
        if i == j {
            return Some(0.0);
        }
        if !self.built {
            return None;
        }
        let mut out = 0f32;
        let rc = unsafe { ffi::mh_library_get_similarity(self.lib, i as i32, j as i32, &mut out) };
        if rc == ffi::MH_OK {
            Some(out)
        } else {
            None
        }
    }

    /// The reference's packed upper triangle, position
    /// i*n - i*(i+1)/2 + j - i - 1 (src/similarity.rs:100), for saving.
    pub fn similarities(&self) -> Result<Vec<f32>> {
        let n = self.lab_colors.len();
        let mut out = vec![0f32; n * n.saturating_sub(1) / 2];
        if self.built && !out.is_empty() {
            ffi::check(unsafe { ffi::mh_library_similarities(self.lib, out.as_mut_ptr(), 1) })?;
        }
        Ok(out)
    }

    /// src/similarity.rs:106-109.
    pub fn get_lab_color(&self, path: &Path) -> Option<Lab> {
        let i = *self.path_to_index.get(path)?;
        self.lab_colors.get(i).copied()
    }

    pub fn len(&self) -> usize {
        self.lab_colors.len()
    }

    pub fn is_empty(&self) -> bool {
        self.lab_colors.is_empty()
    }

    pub fn index_of(&self, path: &Path) -> Option<usize> {
        self.path_to_index.get(path).copied()
    }

    pub fn path_of(&self, index: usize) -> Option<&PathBuf> {
        self.index_to_path.get(&index)
    }

    /// The device library (uploaded on first use).
    fn device(&self) -> Result<*mut ffi::mh_library> {
        if self.lib.is_null() {
            return Err(Error { code: ffi::MH_ERR_INVALID_ARGUMENT,
                               message: "call build_similarities first".into() });
        }
        Ok(self.lib)
    }

    fn grid_indices(&self, grid: &[Vec<Option<PathBuf>>]) -> (Vec<i32>, i32, i32) {
        let h = grid.len();
        let w = if h > 0 { grid[0].len() } else { 0 };
        let mut out = Vec::with_capacity(w * h);
        for row in grid {
            for cell in row {
                out.push(match cell {
                    Some(p) => self.path_to_index.get(p).map(|&i| i as i32).unwrap_or(ffi::MH_EMPTY),
                    None => ffi::MH_EMPTY,
                });
            }
        }
        (out, w as i32, h as i32)
    }

    fn grid_paths(&self, idx: &[i32], grid: &mut [Vec<Option<PathBuf>>]) {
        let w = if grid.is_empty() { 0 } else { grid[0].len() };
        for (y, row) in grid.iter_mut().enumerate() {
            for (x, cell) in row.iter_mut().enumerate() {
                let m = idx[y * w + x];
                *cell = if m < 0 { None } else { self.index_to_path.get(&(m as usize)).cloned() };
            }
        }
    }
}

impl Drop for SimilarityDatabase {
    fn drop(&mut self) {
        if !self.lib.is_null() {
            unsafe { ffi::mh_library_destroy(self.lib) };
        }
    }
}

// ------------------------------------------------------------- adjacency
/// `GridPosition` (src/adjacency.rs:6-15).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub struct GridPosition {
    pub x: usize,
    pub y: usize,
}

impl GridPosition {
    pub fn new(x: usize, y: usize) -> Self {
        GridPosition { x, y }
    }
}

/// `AdjacencyPenaltyCalculator` (src/adjacency.rs:49-214).
pub struct AdjacencyPenaltyCalculator<'a> {
    similarity_db: &'a SimilarityDatabase,
    penalty_weight: f32,
}

impl<'a> AdjacencyPenaltyCalculator<'a> {
    pub fn new(similarity_db: &'a SimilarityDatabase, penalty_weight: f32) -> Self {
        AdjacencyPenaltyCalculator { similarity_db, penalty_weight }
    }

    /// src/adjacency.rs:63-88: weight * sum of 1/(sim+1) over the placed
    /// 4-neighbours (Up, Down, Left, Right).
    pub fn calculate_penalty(&self, candidate_path: &Path, position: GridPosition,
                             grid: &[Vec<Option<PathBuf>>], grid_width: usize,
                             grid_height: usize) -> Result<f32> {
        let db = self.similarity_db;
        let m = match db.index_of(candidate_path) {
            Some(m) => m as i32,
            None => return Ok(0.0),
        };
        let (g, _, _) = db.grid_indices(grid);
        let q = [m, position.x as i32, position.y as i32];
        let mut out = 0f32;
        ffi::check(unsafe {
            ffi::mh_penalties(db.ctx.raw, db.device()?, g.as_ptr(), grid_width as i32,
                              grid_height as i32, self.penalty_weight, q.as_ptr(), 1, &mut out)
        })?;
        Ok(out)
    }

    /// src/adjacency.rs:91-130 (unweighted).
    pub fn calculate_total_cost(&self, grid: &[Vec<Option<PathBuf>>]) -> Result<f32> {
        let db = self.similarity_db;
        let (g, w, h) = db.grid_indices(grid);
        if g.is_empty() {
            return Ok(0.0);
        }
        let mut out = 0f32;
        ffi::check(unsafe { ffi::mh_total_cost(db.ctx.raw, db.device()?, g.as_ptr(), w, h, &mut out) })?;
        Ok(out)
    }

    /// src/adjacency.rs:133-214 (unweighted).
    pub fn calculate_swap_delta(&self, grid: &[Vec<Option<PathBuf>>], pos1: GridPosition,
                                pos2: GridPosition) -> Result<f32> {
        let db = self.similarity_db;
        let (g, w, h) = db.grid_indices(grid);
        if g.is_empty() {
            return Ok(0.0);
        }
        let pair = [pos1.x as i32, pos1.y as i32, pos2.x as i32, pos2.y as i32];
        let mut out = 0f32;
        ffi::check(unsafe {
            ffi::mh_swap_deltas(db.ctx.raw, db.device()?, g.as_ptr(), w, h, pair.as_ptr(), 1,
                                &mut out)
        })?;
        Ok(out)
    }
}

// ------------------------------------------------------------- optimizer
/// `OptimizationConfig` (src/optimizer.rs:7-27); `seed` keys the counter RNG
/// that replaces the unseeded thread_rng.
#[derive(Clone, Debug)]
pub struct OptimizationConfig {
    pub max_iterations: usize,
    pub initial_temperature: f32,
    pub temperature_decay: f32,
    pub report_interval: usize,
    pub seed: u64,
}

impl Default for OptimizationConfig {
    fn default() -> Self {
        OptimizationConfig { max_iterations: 1000, initial_temperature: 100.0,
                             temperature_decay: 0.99995, report_interval: 100, seed: 0 }
    }
}

/// `OptimizationResult` (src/optimizer.rs:202-221).
#[derive(Clone, Debug, Default)]
pub struct OptimizationResult {
    pub initial_cost: f32,
    pub final_cost: f32,
    pub best_cost: f32,
    pub improved_count: usize,
    pub accepted_count: usize,
    pub iterations: usize,
}

impl OptimizationResult {
    pub fn improvement_percentage(&self) -> f32 {
        if self.initial_cost > 0.0 {
            ((self.initial_cost - self.final_cost) / self.initial_cost) * 100.0
        } else {
            0.0
        }
    }

    fn from_raw(r: &ffi::mh_anneal_result) -> Self {
        OptimizationResult { initial_cost: r.initial_cost, final_cost: r.final_cost,
                             best_cost: r.best_cost, improved_count: r.improved_count as usize,
                             accepted_count: r.accepted_count as usize,
                             iterations: r.iterations as usize }
    }
}

/// `MosaicOptimizer` (src/optimizer.rs:30-198).
pub struct MosaicOptimizer<'a> {
    calculator: &'a AdjacencyPenaltyCalculator<'a>,
    config: OptimizationConfig,
}

impl<'a> MosaicOptimizer<'a> {
    pub fn new(calculator: &'a AdjacencyPenaltyCalculator<'a>, config: OptimizationConfig) -> Self {
        MosaicOptimizer { calculator, config }
    }

    fn run(&self, grid: &mut [Vec<Option<PathBuf>>], iterations: usize, greedy: bool)
           -> Result<OptimizationResult> {
        let db = self.calculator.similarity_db;
        let (mut g, w, h) = db.grid_indices(grid);
        if g.is_empty() {
            return Ok(OptimizationResult::default());
        }
        let p = ffi::mh_anneal_params { max_iterations: iterations as i64,
                                        initial_temperature: self.config.initial_temperature,
                                        temperature_decay: self.config.temperature_decay,
                                        seed: self.config.seed, greedy: greedy as i32 };
        let mut r = ffi::mh_anneal_result::default();
        ffi::check(unsafe {
            ffi::mh_anneal(db.ctx.raw, db.device()?, g.as_mut_ptr(), w, h, &p, &mut r,
                           ptr::null_mut(), ptr::null_mut())
        })?;
        db.grid_paths(&g, grid);
        Ok(OptimizationResult::from_raw(&r))
    }

    /// Simulated annealing (src/optimizer.rs:41-129), the grid swapped in place.
    pub fn optimize(&self, grid: &mut [Vec<Option<PathBuf>>]) -> Result<OptimizationResult> {
        self.run(grid, self.config.max_iterations, false)
    }

    /// Improvements only (src/optimizer.rs:132-198).
    pub fn optimize_greedy(&self, grid: &mut [Vec<Option<PathBuf>>], max_iterations: usize)
                           -> Result<OptimizationResult> {
        self.run(grid, max_iterations, true)
    }
}

// ------------------------------------------------------ colour adjustment
/// `ColorAdjustment` (src/color_adjustment.rs:6-96).
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct ColorAdjustment {
    pub brightness: f32,
    pub contrast: f32,
    pub hue_shift: f32,
    pub saturation: f32,
}

impl Default for ColorAdjustment {
    fn default() -> Self {
        ColorAdjustment { brightness: 0.0, contrast: 1.0, hue_shift: 0.0, saturation: 1.0 }
    }
}

impl ColorAdjustment {
    /// src/color_adjustment.rs:30-37: each factor clamped.
    pub fn new(brightness: f32, contrast: f32, hue_shift: f32, saturation: f32) -> Self {
        ColorAdjustment { brightness: brightness.clamp(-1.0, 1.0), contrast: contrast.clamp(0.0, 2.0),
                          hue_shift: hue_shift.clamp(-180.0, 180.0),
                          saturation: saturation.clamp(0.0, 2.0) }
    }

    fn as_array(&self) -> [f32; 4] {
        [self.brightness, self.contrast, self.hue_shift, self.saturation]
    }

    /// src/color_adjustment.rs:40-50 over RGB8 pixels.
    pub fn apply_to_image(&self, ctx: &Context, rgb: &[u8]) -> Result<Vec<u8>> {
        let mut out = vec![0u8; rgb.len()];
        let a = self.as_array();
        ffi::check(unsafe {
            ffi::mh_adjust_pixels(ctx.raw, a.as_ptr(), rgb.as_ptr(), (rgb.len() / 3) as i64,
                                  out.as_mut_ptr())
        })?;
        Ok(out)
    }

    /// src/color_adjustment.rs:53-87.
    pub fn adjust_pixel(&self, ctx: &Context, pixel: [u8; 3]) -> Result<[u8; 3]> {
        let v = self.apply_to_image(ctx, &pixel)?;
        Ok([v[0], v[1], v[2]])
    }
}

/// src/color_adjustment.rs:99-154.
pub fn calculate_optimal_adjustment(ctx: &Context, tile_avg_rgb: [u8; 3], target_avg_rgb: [u8; 3],
                                    adjustment_strength: f32) -> Result<ColorAdjustment> {
    let mut out = [0f32; 4];
    ffi::check(unsafe {
        ffi::mh_optimal_adjustments(ctx.raw, tile_avg_rgb.as_ptr(), target_avg_rgb.as_ptr(), 1,
                                    adjustment_strength, out.as_mut_ptr())
    })?;
    Ok(ColorAdjustment { brightness: out[0], contrast: out[1], hue_shift: out[2],
                         saturation: out[3] })
}

// --------------------------------------------------------- whole mosaic
/// Material images resident on the GPU (pixels and Lab means), in the
/// order of `paths`: `load_tiles` / `process_tile` (src/main.rs:166-301)
/// minus the file decode.
pub struct MaterialLibrary {
    ctx: Rc<Context>,
    paths: Vec<PathBuf>,
    raw: *mut ffi::mh_library,
}

impl MaterialLibrary {
    /// images[i] = (rgb, width, height) of paths[i]; any sizes.
    pub fn from_images(ctx: Rc<Context>, paths: Vec<PathBuf>, images: &[(&[u8], u32, u32)])
                       -> Result<Self> {
        assert_eq!(paths.len(), images.len());
        let mut flat = Vec::new();
        let mut offsets = Vec::with_capacity(images.len());
        let mut widths = Vec::with_capacity(images.len());
        let mut heights = Vec::with_capacity(images.len());
        for (rgb, w, h) in images {
            offsets.push(flat.len() as i64);
            widths.push(*w as i32);
            heights.push(*h as i32);
            flat.extend_from_slice(rgb);
        }
        let mut raw = ptr::null_mut();
        ffi::check(unsafe {
            ffi::mh_library_from_images(ctx.raw, flat.as_ptr(), offsets.as_ptr(), widths.as_ptr(),
                                        heights.as_ptr(), images.len() as i32, &mut raw)
        })?;
        Ok(MaterialLibrary { ctx, paths, raw })
    }

    pub fn paths(&self) -> &[PathBuf] {
        &self.paths
    }
}

impl Drop for MaterialLibrary {
    fn drop(&mut self) {
        unsafe { ffi::mh_library_destroy(self.raw) };
    }
}

/// The generator's settings (src/main.rs:477-490).
#[derive(Clone, Debug)]
pub struct MosaicSettings {
    pub grid_w: u32,
    pub grid_h: u32,
    pub max_usage_per_image: usize,
    pub adjacency_penalty_weight: f32,
    pub enable_optimization: bool,
    pub optimization_iterations: usize,
    pub color_adjustment_strength: f32,
    pub seed: u64,
}

/// A placed, optionally annealed and composited mosaic.
pub struct Mosaic {
    pub grid: Vec<Vec<Option<PathBuf>>>,
    pub image: Vec<u8>,
    pub width: u32,
    pub height: u32,
    pub optimization: Option<OptimizationResult>,
}

/// generate_mosaic (src/main.rs:477-685) minus file I/O: cell Lab means, the
/// K nearest materials, the placement chain, simulated annealing and the
/// colour-adjusted composite, on a target already decoded and resized.
pub fn generate_mosaic(lib: &MaterialLibrary, target: &[u8], width: u32, height: u32,
                       s: &MosaicSettings) -> Result<Mosaic> {
    let (gw, gh) = (s.grid_w as usize, s.grid_h as usize);
    let (tw, th) = (width as usize / gw.max(1), height as usize / gh.max(1));
    let params = ffi::mh_job_params {
        grid_w: s.grid_w as i32, grid_h: s.grid_h as i32,
        max_usage: s.max_usage_per_image as i32, weight: s.adjacency_penalty_weight,
        color_strength: s.color_adjustment_strength.clamp(0.0, 1.0),
        enable_optimization: s.enable_optimization as i32,
        optimization_iterations: s.optimization_iterations as i64,
        initial_temperature: 100.0, temperature_decay: 0.99995, seed: s.seed,
        score_mode: ffi::MH_SCORE_SQDIST,
    };
    let mut grid = vec![ffi::MH_EMPTY; gw * gh];
    let mut image = vec![0u8; gw * tw * gh * th * 3];
    let mut res = ffi::mh_job_result {
        place: ffi::mh_place_stats::default(),
        anneal: ffi::mh_anneal_result::default(),
        out_width: 0,
        out_height: 0,
    };
    ffi::check(unsafe {
        ffi::mh_generate(lib.ctx.raw, lib.raw, target.as_ptr(), width as i32, height as i32,
                         &params, grid.as_mut_ptr(), image.as_mut_ptr(), &mut res)
    })?;
    let rows = (0..gh)
        .map(|y| {
            (0..gw)
                .map(|x| {
                    let m = grid[y * gw + x];
                    if m < 0 { None } else { lib.paths.get(m as usize).cloned() }
                })
                .collect()
        })
        .collect();
    let ran = s.enable_optimization && s.adjacency_penalty_weight > 0.0;
    Ok(Mosaic { grid: rows, image, width: res.out_width as u32, height: res.out_height as u32,
                optimization: if ran { Some(OptimizationResult::from_raw(&res.anneal)) } else { None } })
}
