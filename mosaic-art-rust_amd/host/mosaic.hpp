// mosaic.hpp -- C++ host API of the gfx950 tile-matching path, mirroring the
// public Rust API of mosaic-art-rust over the C ABI (include/mosaic_hip.h).
//
// Names, argument meaning and error behaviour follow the reference:
//   SimilarityDatabase            src/similarity.rs:30-137
//   GridPosition                  src/adjacency.rs:5-46
//   AdjacencyPenaltyCalculator    src/adjacency.rs:48-214
//   OptimizationConfig / Result   src/optimizer.rs:6-27, :202-221
//   MosaicOptimizer               src/optimizer.rs:29-198
//   ColorAdjustment               src/color_adjustment.rs:5-96
//   calculate_optimal_adjustment  src/color_adjustment.rs:99-154
//   UsageTracker                  src/lib.rs:57-89
//   MosaicGeneratorImpl           src/lib.rs:91-124
//   Tile, load_tiles, process_tile src/lib.rs:44-50, src/main.rs:166-301
//   similarity-DB preparation     src/main.rs:119-143 (MosaicGenerator::new)
// Differences a caller sees:
//   * paths are std::string, grids are Grid (grid[y][x], std::nullopt = None);
//   * Option<T> is std::optional<T>; anyhow::Result errors are thrown as
//     mosaic::Error (with the C ABI status code);
//   * the annealing random stream is seeded (OptimizationConfig::seed) instead
//     of thread_rng, so runs are reproducible;
//   * `new` constructors are C++ constructors, except ColorAdjustment::create
//     (the clamping constructor; the default constructor is Default);
//   * images are RgbImage (packed RGB8 rows) instead of image::DynamicImage.
// Computation runs on the GPU through the C ABI; there is no CPU fallback
// (a missing device raises mosaic::Error with MH_ERR_NO_DEVICE).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mosaic_hip.h"

namespace mosaic {

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

// One HIP device and stream (mh_context).  default_context() is created on
// first use and shared by objects built without an explicit context.
class Context {
   public:
    explicit Context(int32_t device = 0);
    ~Context();
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    mh_context* get() const { return h_; }
    static std::shared_ptr<Context> default_context();

   private:
    mh_context* h_ = nullptr;
};

struct Lab {  // palette::Lab<D65, f32>
    float l = 0.0f, a = 0.0f, b = 0.0f;
};

struct Rgb {  // image::Rgb<u8>
    uint8_t r = 0, g = 0, b = 0;
    bool operator==(const Rgb& o) const { return r == o.r && g == o.g && b == o.b; }
};

struct RgbImage {  // image::RgbImage: width x height, packed RGB8 rows
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> data;
    RgbImage() = default;
    RgbImage(uint32_t w, uint32_t h, Rgb fill = Rgb{});
    Rgb get_pixel(uint32_t x, uint32_t y) const;
    void put_pixel(uint32_t x, uint32_t y, Rgb p);
};

using Path = std::string;
using Grid = std::vector<std::vector<std::optional<Path>>>;  // grid[y][x]

// ------------------------------------------------------------ similarity
class SimilarityDatabase {
   public:
    SimilarityDatabase();  // SimilarityDatabase::new / Default
    explicit SimilarityDatabase(std::shared_ptr<Context> ctx);
    ~SimilarityDatabase();
    SimilarityDatabase(const SimilarityDatabase& o);
    SimilarityDatabase& operator=(const SimilarityDatabase& o);
    SimilarityDatabase(SimilarityDatabase&& o) noexcept;
    SimilarityDatabase& operator=(SimilarityDatabase&& o) noexcept;

    void add_tile(const Path& path, const Lab& lab_color);
    // All pairwise Lab distances (GPU), kept as the packed upper triangle.
    void build_similarities();
    // Distance between two tiles; 0.0 for the same tile; nullopt for an
    // unknown path or before build_similarities.
    std::optional<float> get_similarity(const Path& path1, const Path& path2) const;
    std::optional<Lab> get_lab_color(const Path& path) const;
    // serde_json pretty format of the reference database (src/similarity.rs:112-137).
    void save_to_file(const Path& path) const;
    static SimilarityDatabase load_from_file(const Path& path);
    static SimilarityDatabase load_or_new(const Path& path);

    size_t size() const { return lab_colors_.size(); }
    const std::vector<Lab>& lab_colors() const { return lab_colors_; }
    const std::vector<float>& similarities() const { return similarities_; }
    std::optional<int32_t> index_of(const Path& path) const;
    const Path* path_of(int32_t index) const;
    // The device-resident library of the current tiles (built on demand).
    mh_library* device_library() const;
    const std::shared_ptr<Context>& context() const { return ctx_; }

   private:
    std::shared_ptr<Context> ctx_;
    std::unordered_map<Path, size_t> path_to_index_;
    std::map<size_t, Path> index_to_path_;
    std::vector<Lab> lab_colors_;
    std::vector<float> similarities_;
    mutable mh_library* lib_ = nullptr;
    mutable size_t lib_n_ = 0;
    void drop_library();
};

// ------------------------------------------------------------- adjacency
struct GridPosition {
    size_t x = 0, y = 0;
    GridPosition() = default;
    GridPosition(size_t x, size_t y) : x(x), y(y) {}
    bool operator==(const GridPosition& o) const { return x == o.x && y == o.y; }
    // Up, Down, Left, Right (src/adjacency.rs:17-45).
    std::vector<GridPosition> get_adjacent_positions(size_t grid_width, size_t grid_height) const;
};

class AdjacencyPenaltyCalculator {
   public:
    AdjacencyPenaltyCalculator(const SimilarityDatabase& similarity_db, float penalty_weight);
    // Sum of 1/(sim+1) over the placed neighbours, times the weight.
    float calculate_penalty(const Path& candidate_path, GridPosition position, const Grid& grid,
                            size_t grid_width, size_t grid_height) const;
    // Sum of 1/(sim+1) over right and down neighbour pairs (unweighted).
    float calculate_total_cost(const Grid& grid) const;
    // Change of the total cost if the tiles at pos1 and pos2 are swapped.
    float calculate_swap_delta(const Grid& grid, GridPosition pos1, GridPosition pos2) const;
    const SimilarityDatabase& similarity_db() const { return *db_; }

   private:
    const SimilarityDatabase* db_;
    float penalty_weight_;
};

// ------------------------------------------------------------- optimizer
struct OptimizationConfig {
    size_t max_iterations = 1000;
    float initial_temperature = 100.0f;
    float temperature_decay = 0.99995f;
    size_t report_interval = 100;  // kept for the reference's field set; nothing is printed
    uint64_t seed = 0;             // counter-RNG key (the reference uses thread_rng)
};

struct OptimizationResult {
    float initial_cost = 0.0f, final_cost = 0.0f, best_cost = 0.0f;
    size_t improved_count = 0, accepted_count = 0, iterations = 0;
    float improvement_percentage() const;
};

class MosaicOptimizer {
   public:
    MosaicOptimizer(const AdjacencyPenaltyCalculator& calculator, OptimizationConfig config);
    OptimizationResult optimize(Grid& grid) const;
    OptimizationResult optimize_greedy(Grid& grid, size_t max_iterations) const;

   private:
    const AdjacencyPenaltyCalculator* calc_;
    OptimizationConfig config_;
    OptimizationResult run(Grid& grid, size_t iterations, bool greedy) const;
};

// ------------------------------------------------------ colour adjustment
struct ColorAdjustment {
    float brightness = 0.0f;  // -1.0 .. 1.0
    float contrast = 1.0f;    // 0.0 .. 2.0
    float hue_shift = 0.0f;   // -180 .. 180 degrees
    float saturation = 1.0f;  // 0.0 .. 2.0
    // ColorAdjustment::new: the clamping constructor.
    static ColorAdjustment create(float brightness, float contrast, float hue_shift,
                                  float saturation);
    RgbImage apply_to_image(const RgbImage& img) const;
    Rgb adjust_pixel(Rgb pixel) const;
};

ColorAdjustment calculate_optimal_adjustment(Rgb tile_avg_rgb, Rgb target_avg_rgb,
                                             float adjustment_strength);

// ------------------------------------------------------------------ lib
class UsageTracker {
   public:
    explicit UsageTracker(size_t max_usage_per_image) : max_(max_usage_per_image) {}
    bool can_use_image(const Path& path) const;
    void use_image(const Path& path);
    size_t get_usage_count(const Path& path) const;
    void reset() { counts_.clear(); }

   private:
    std::unordered_map<Path, size_t> counts_;
    size_t max_;
};

struct MosaicGeneratorImpl {
    // Average Lab of an image: serial f32 fold of per-pixel Lab in row-major
    // order, divided by the pixel count (GPU).
    static Lab calculate_average_lab(const RgbImage& img);
    static bool is_aspect_ratio_match(float img_aspect, float target_aspect, float tolerance);
};

// ------------------------------------------------------- material ingest
struct Tile {  // src/lib.rs:44-50
    Path path;
    Lab lab_color;
    float aspect_ratio = 0.0f;  // width / height in f32
    size_t source_index = 0;    // position of the image in load_tiles' entries
};

// One directory entry handed to load_tiles: the path and the decoded RGB8
// image, or std::nullopt when the file did not decode (the reference logs
// such files and skips them, src/main.rs:200-206).  Decoding and directory
// scanning stay with the caller.
struct MaterialSource {
    Path path;
    std::optional<RgbImage> image;
};

// The extension filter of the directory scan (src/main.rs:172-184): png,
// jpg, jpeg or webp in any letter case.
bool is_material_file(const Path& path);

// The entries load_tiles keeps, in entry order (src/main.rs:166-261): those
// whose aspect ratio matches; when none does, the decodable ones among the
// first min(entries, 2 * max_materials); then at most max_materials of them.
// sizes[i] = {width, height}, or std::nullopt for an entry that failed.
struct MaterialSelection {
    std::vector<size_t> picked;
    bool fallback = false;  // no aspect match: the unfiltered sample was taken
};
MaterialSelection select_materials(
    const std::vector<std::optional<std::pair<uint32_t, uint32_t>>>& sizes, float target_aspect,
    float aspect_tolerance, size_t max_materials);

// load_tiles over decoded entries: select_materials, then the average Lab of
// every picked image from one batched GPU launch (mh_average_lab_many).
std::vector<Tile> load_tiles(const std::vector<MaterialSource>& entries, float target_aspect,
                             float aspect_tolerance, size_t max_materials,
                             std::shared_ptr<Context> ctx = nullptr);

// MosaicGenerator::new's database step (src/main.rs:119-143): a fresh
// database of the tiles, built and saved, when rebuild_similarity is set or
// db_path does not exist (a failed save only warns on stderr); otherwise
// load_or_new(db_path).  Tiles missing from it are then added and the
// similarities rebuilt.
SimilarityDatabase prepare_similarity_db(const std::vector<Tile>& tiles, const Path& db_path,
                                         bool rebuild_similarity);

// ----------------------------------------------------------- whole job
// generate_mosaic minus file I/O: materials are RGB images of any sizes
// (material i is tile i of the result grid's indices; each is resized from
// its own size to the cell size), the target is resized by the caller to the
// output size (src/main.rs:491-505).
// Candidate score of the placement scan: the CLI's squared Lab distance or
// the GUI generator's Euclidean distance (src/gui/app_full.rs:1436-1448),
// plus the adjacency penalty in both.
enum class ScoreMode : int32_t { SquaredDistance = MH_SCORE_SQDIST, Distance = MH_SCORE_DISTANCE };

struct MosaicSettings {
    int32_t grid_w = 50, grid_h = 28;
    int32_t max_usage_per_image = 3;
    float adjacency_penalty_weight = 0.3f;
    float color_adjustment_strength = 0.3f;
    bool enable_optimization = true;
    size_t optimization_iterations = 1000;
    OptimizationConfig optimization;
    ScoreMode score_mode = ScoreMode::SquaredDistance;
};

struct MosaicOutput {
    RgbImage image;
    std::vector<int32_t> grid;  // grid_w * grid_h material indices, -1 = empty
    mh_job_result result{};
};

class MaterialLibrary {  // device-resident materials (pixels + Lab table)
   public:
    explicit MaterialLibrary(const std::vector<RgbImage>& materials,
                             std::shared_ptr<Context> ctx = nullptr);
    ~MaterialLibrary();
    MaterialLibrary(const MaterialLibrary&) = delete;
    MaterialLibrary& operator=(const MaterialLibrary&) = delete;
    MosaicOutput generate_mosaic(const RgbImage& target, const MosaicSettings& settings) const;
    int32_t size() const;
    mh_library* get() const { return lib_; }

   private:
    std::shared_ptr<Context> ctx_;
    mh_library* lib_ = nullptr;
};

}  // namespace mosaic
