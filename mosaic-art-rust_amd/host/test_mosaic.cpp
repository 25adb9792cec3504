// test_mosaic.cpp -- the reference's unit tests, restated against the C++
// host API (test names follow src/similarity.rs, src/adjacency.rs,
// src/optimizer.rs, src/color_adjustment.rs and src/lib.rs).  Tests marked
// HOST run without a GPU (bookkeeping and the database file format); the rest
// call the gfx950 kernels through the C ABI.
//
//   test_mosaic            run every test (needs a GPU)
//   test_mosaic --host     run the HOST tests only
//   test_mosaic --list     print the test names
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <unistd.h>
#include <vector>

#include "mosaic.hpp"

using namespace mosaic;

namespace {

struct TestCase {
    const char* name;
    bool host_only;
    std::function<void()> fn;
};

std::vector<TestCase>& registry() {
    static std::vector<TestCase> r;
    return r;
}

struct Reg {
    Reg(const char* n, bool h, std::function<void()> f) { registry().push_back({n, h, std::move(f)}); }
};

struct Failure {
    std::string msg;
};

#define TEST(name) \
    static void name(); \
    static Reg reg_##name(#name, false, name); \
    static void name()
#define HOST_TEST(name) \
    static void name(); \
    static Reg reg_##name(#name, true, name); \
    static void name()
#define CHECK(cond)                                                                        \
    do {                                                                                   \
        if (!(cond)) throw Failure{std::string(__FILE__) + ":" + std::to_string(__LINE__) + \
                                   ": " #cond};                                            \
    } while (0)
#define CHECK_EQ(a, b) CHECK((a) == (b))
#define CHECK_NEAR(a, b, tol) CHECK(std::fabs((double)(a) - (double)(b)) <= (tol))

std::string temp_path(const char* tag) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "/tmp/mosaic_%s_%d.json", tag, (int)getpid());
    return buf;
}

Lab lab(float l, float a, float b) { return Lab{l, a, b}; }

// src/optimizer.rs:236-253
std::pair<Grid, SimilarityDatabase> create_test_grid() {
    SimilarityDatabase db;
    db.add_tile("tile1.png", lab(50, 0, 0));
    db.add_tile("tile2.png", lab(60, 10, 10));
    db.add_tile("tile3.png", lab(40, -10, -10));
    db.add_tile("tile4.png", lab(55, 5, 5));
    db.build_similarities();
    Grid g(2, std::vector<std::optional<Path>>(2));
    g[0][0] = "tile1.png";
    g[0][1] = "tile2.png";
    g[1][0] = "tile3.png";
    g[1][1] = "tile4.png";
    return {g, db};
}

}  // namespace

// ============================================================ similarity.rs
TEST(test_similarity_database_basic) {
    SimilarityDatabase db;
    db.add_tile("tile1.png", lab(50, 0, 0));
    db.add_tile("tile2.png", lab(60, 10, 10));
    db.add_tile("tile3.png", lab(40, -10, -10));
    db.build_similarities();
    CHECK(db.get_similarity("tile1.png", "tile2.png").has_value());
    CHECK(!db.get_lab_color("nonexistent.png").has_value());
}

TEST(test_lab_distance_calculation) {
    // (50,0,0) vs (60,10,10): sqrt(100 + 100 + 100)
    SimilarityDatabase db;
    db.add_tile("a.png", lab(50, 0, 0));
    db.add_tile("b.png", lab(60, 10, 10));
    db.build_similarities();
    CHECK_NEAR(*db.get_similarity("a.png", "b.png"), std::sqrt(300.0f), 1e-4);
}

HOST_TEST(test_similarity_database_get_lab_color) {
    SimilarityDatabase db;
    db.add_tile("red.png", lab(53.24f, 80.09f, 67.2f));
    auto c = db.get_lab_color("red.png");
    CHECK(c.has_value());
    CHECK_EQ(c->l, 53.24f);
    CHECK_EQ(c->a, 80.09f);
    CHECK_EQ(c->b, 67.2f);
    CHECK(!db.get_lab_color("missing.png").has_value());
}

TEST(test_similarity_database_save_load) {
    SimilarityDatabase db;
    db.add_tile("tile1.png", lab(50, 0, 0));
    db.add_tile("tile2.png", lab(60, 10, 10));
    db.build_similarities();
    const std::string p = temp_path("save_load");
    db.save_to_file(p);
    SimilarityDatabase loaded = SimilarityDatabase::load_from_file(p);
    CHECK_EQ(loaded.lab_colors().size(), 2u);
    auto sim = loaded.get_similarity("tile1.png", "tile2.png");
    CHECK(sim.has_value());
    CHECK(*sim > 0.0f);
    CHECK_EQ(*sim, *db.get_similarity("tile1.png", "tile2.png"));
    std::remove(p.c_str());
}

HOST_TEST(test_similarity_database_load_or_new) {
    SimilarityDatabase db = SimilarityDatabase::load_or_new("/tmp/nonexistent_db_for_mosaic.json");
    CHECK_EQ(db.lab_colors().size(), 0u);
    SimilarityDatabase original;
    original.add_tile("test.png", lab(50, 0, 0));
    const std::string p = temp_path("load_or_new");
    original.save_to_file(p);
    SimilarityDatabase loaded = SimilarityDatabase::load_or_new(p);
    CHECK_EQ(loaded.lab_colors().size(), 1u);
    std::remove(p.c_str());
}

HOST_TEST(test_similarity_database_load_from_invalid_file) {
    const std::string p = temp_path("invalid");
    {
        std::ofstream f(p);
        f << "invalid json content";
    }
    bool threw = false;
    try {
        SimilarityDatabase::load_from_file(p);
    } catch (const Error&) {
        threw = true;
    }
    CHECK(threw);
    std::remove(p.c_str());
}

TEST(test_similarity_database_empty) {
    SimilarityDatabase db;
    db.build_similarities();  // no tiles: nothing to build
    CHECK(!db.get_similarity("a.png", "b.png").has_value());
    CHECK(!db.get_lab_color("a.png").has_value());
}

TEST(test_similarity_database_single_tile) {
    SimilarityDatabase db;
    db.add_tile("single.png", lab(50, 0, 0));
    db.build_similarities();
    CHECK_EQ(*db.get_similarity("single.png", "single.png"), 0.0f);
    CHECK(db.get_lab_color("single.png").has_value());
}

TEST(test_similarity_database_multiple_tiles) {
    SimilarityDatabase db;
    const char* names[] = {"t0.png", "t1.png", "t2.png", "t3.png", "t4.png"};
    for (int i = 0; i < 5; ++i) db.add_tile(names[i], lab(10.0f * i, (float)i, -(float)i));
    db.build_similarities();
    CHECK_EQ(db.similarities().size(), 10u);
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            auto s1 = db.get_similarity(names[i], names[j]);
            auto s2 = db.get_similarity(names[j], names[i]);
            CHECK(s1.has_value() && s2.has_value());
            CHECK_EQ(*s1, *s2);  // symmetric
            if (i == j) CHECK_EQ(*s1, 0.0f);
            else CHECK(*s1 > 0.0f);
        }
}

HOST_TEST(test_database_default) {
    SimilarityDatabase db;
    CHECK_EQ(db.size(), 0u);
    CHECK(db.similarities().empty());
}

// The reference writes serde_json pretty output; a file in that layout (key
// order differs run to run there) must load, and a saved file must load back
// bit for bit.
HOST_TEST(test_similarity_database_reference_json_layout) {
    const std::string p = temp_path("layout");
    {
        std::ofstream f(p);
        f << "{\n  \"path_to_index\": {\n    \"tiles/b.png\": 1,\n    \"tiles/a.png\": 0\n  },\n"
             "  \"index_to_path\": {\n    \"1\": \"tiles/b.png\",\n    \"0\": \"tiles/a.png\"\n  },\n"
             "  \"lab_colors\": [\n    {\n      \"l\": 50.0,\n      \"a\": 0.0,\n      \"b\": -1.5e-7\n"
             "    },\n    {\n      \"l\": 60.0,\n      \"a\": 10.0,\n      \"b\": 10.0\n    }\n  ],\n"
             "  \"similarities\": [\n    17.320509\n  ]\n}";
    }
    SimilarityDatabase db = SimilarityDatabase::load_from_file(p);
    CHECK_EQ(db.size(), 2u);
    CHECK_EQ(*db.get_similarity("tiles/a.png", "tiles/b.png"), 17.320509f);
    CHECK_EQ(db.get_lab_color("tiles/a.png")->b, -1.5e-7f);
    db.save_to_file(p);
    SimilarityDatabase again = SimilarityDatabase::load_from_file(p);
    CHECK_EQ(again.get_lab_color("tiles/a.png")->b, -1.5e-7f);
    CHECK_EQ(*again.get_similarity("tiles/b.png", "tiles/a.png"), 17.320509f);
    std::ifstream f(p);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    CHECK(text.find("\"l\": 50.0,") != std::string::npos);
    CHECK(text.find("\"b\": -1.5e-7") != std::string::npos);
    CHECK(text.find("    17.320509\n") != std::string::npos);
    std::remove(p.c_str());
}

// ============================================================= adjacency.rs
HOST_TEST(test_grid_position_adjacent) {
    GridPosition pos(1, 1);
    auto adj = pos.get_adjacent_positions(3, 3);
    CHECK_EQ(adj.size(), 4u);
    auto has = [&](const std::vector<GridPosition>& v, GridPosition p) {
        for (auto& q : v)
            if (q == p) return true;
        return false;
    };
    CHECK(has(adj, GridPosition(1, 0)));
    CHECK(has(adj, GridPosition(1, 2)));
    CHECK(has(adj, GridPosition(0, 1)));
    CHECK(has(adj, GridPosition(2, 1)));
    auto corner = GridPosition(0, 0).get_adjacent_positions(3, 3);
    CHECK_EQ(corner.size(), 2u);
    CHECK(has(corner, GridPosition(1, 0)));
    CHECK(has(corner, GridPosition(0, 1)));
}

TEST(test_adjacency_penalty_calculation) {
    SimilarityDatabase db;
    db.add_tile("tile1.png", lab(50, 0, 0));
    db.add_tile("tile2.png", lab(60, 10, 10));
    db.add_tile("tile3.png", lab(40, -10, -10));
    db.build_similarities();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    Grid grid(3, std::vector<std::optional<Path>>(3));
    grid[0][0] = "tile1.png";
    grid[0][1] = "tile2.png";
    const float penalty = calc.calculate_penalty("tile3.png", GridPosition(1, 0), grid, 3, 3);
    CHECK(penalty > 0.0f);
    // the single placed neighbour (tile1 at (0,0)) gives exactly 1/(d+1)
    const float d = *db.get_similarity("tile3.png", "tile1.png");
    CHECK_EQ(penalty, 1.0f / (d + 1.0f));
}

TEST(test_total_cost_calculation) {
    SimilarityDatabase db;
    db.add_tile("tile1.png", lab(50, 0, 0));
    db.add_tile("tile2.png", lab(50, 0, 0));  // same colour
    db.build_similarities();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    Grid grid(2, std::vector<std::optional<Path>>(2));
    grid[0][0] = "tile1.png";
    grid[0][1] = "tile2.png";
    const float total = calc.calculate_total_cost(grid);
    CHECK(total > 0.5f);
    CHECK_EQ(total, 1.0f);  // identical colours: similarity 0, one right pair
}

TEST(test_swap_delta_matches_cost_difference) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    const float before = calc.calculate_total_cost(grid);
    const float delta = calc.calculate_swap_delta(grid, GridPosition(0, 0), GridPosition(1, 1));
    std::swap(grid[0][0], grid[1][1]);
    const float after = calc.calculate_total_cost(grid);
    CHECK_NEAR(delta, after - before, 1e-5);
}

TEST(test_unbuilt_similarities_are_refused) {
    SimilarityDatabase db;
    db.add_tile("a.png", lab(50, 0, 0));
    db.add_tile("b.png", lab(60, 0, 0));
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    Grid grid(1, std::vector<std::optional<Path>>(2));
    grid[0][0] = "a.png";
    grid[0][1] = "b.png";
    bool threw = false;
    try {
        calc.calculate_total_cost(grid);
    } catch (const Error& e) {
        threw = e.code == MH_ERR_INVALID_ARGUMENT;
    }
    CHECK(threw);
}

// ============================================================= optimizer.rs
TEST(test_optimization_basic) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 10;
    MosaicOptimizer opt(calc, cfg);
    auto r = opt.optimize(grid);
    CHECK_EQ(r.iterations, 10u);
    CHECK(r.final_cost >= 0.0f);
}

TEST(test_greedy_optimization) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    MosaicOptimizer opt(calc, OptimizationConfig{});
    auto r = opt.optimize_greedy(grid, 10);
    CHECK(r.final_cost <= r.initial_cost);
    CHECK_EQ(r.improved_count, r.accepted_count);
}

HOST_TEST(test_optimization_result_improvement) {
    OptimizationResult r;
    r.initial_cost = 100.0f;
    r.final_cost = 75.0f;
    r.best_cost = 75.0f;
    r.improved_count = 10;
    r.accepted_count = 15;
    r.iterations = 100;
    CHECK_EQ(r.improvement_percentage(), 25.0f);
}

TEST(test_optimization_empty_grid) {
    Grid empty;
    SimilarityDatabase db;
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    MosaicOptimizer opt(calc, OptimizationConfig{});
    auto r = opt.optimize(empty);
    CHECK_EQ(r.initial_cost, 0.0f);
    CHECK_EQ(r.final_cost, 0.0f);
    CHECK_EQ(r.improved_count, 0u);
}

TEST(test_optimization_single_tile_grid) {
    SimilarityDatabase db;
    db.add_tile("single.png", lab(50, 0, 0));
    db.build_similarities();
    Grid grid{{Path("single.png")}};
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 100;
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK_EQ(r.initial_cost, 0.0f);
    CHECK_EQ(r.final_cost, 0.0f);
    CHECK_EQ(r.improved_count, 0u);
    CHECK_EQ(r.accepted_count, 0u);
}

TEST(test_optimization_grid_with_empty_cells) {
    SimilarityDatabase db;
    db.add_tile("tile1.png", lab(50, 0, 0));
    db.add_tile("tile2.png", lab(60, 10, 10));
    db.build_similarities();
    Grid grid(3, std::vector<std::optional<Path>>(3));
    grid[0][0] = "tile1.png";
    grid[2][2] = "tile2.png";
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 50;
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK(r.final_cost >= 0.0f);
    CHECK_EQ(r.initial_cost, 0.0f);
}

TEST(test_optimization_extreme_temperature_values) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig hot;
    hot.max_iterations = 20;
    hot.initial_temperature = 10000.0f;
    hot.temperature_decay = 0.95f;
    auto r1 = MosaicOptimizer(calc, hot).optimize(grid);
    CHECK_EQ(r1.iterations, 20u);
    CHECK(r1.final_cost >= 0.0f);
    auto [grid2, db2] = create_test_grid();
    (void)db2;
    OptimizationConfig cold;
    cold.max_iterations = 20;
    cold.initial_temperature = 0.001f;
    cold.temperature_decay = 0.99f;
    auto r2 = MosaicOptimizer(calc, cold).optimize(grid2);
    CHECK_EQ(r2.iterations, 20u);
    CHECK(r2.final_cost <= r2.initial_cost || std::isfinite(r2.final_cost));
}

TEST(test_optimization_zero_temperature_decay) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 10;
    cfg.temperature_decay = 0.0f;
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK_EQ(r.iterations, 10u);
    CHECK(std::isfinite(r.final_cost));
}

TEST(test_optimization_very_fast_decay) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 50;
    cfg.initial_temperature = 1000.0f;
    cfg.temperature_decay = 0.5f;
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK_EQ(r.iterations, 50u);
    CHECK(r.final_cost >= 0.0f);
}

TEST(test_optimization_zero_iterations) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 0;
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK_EQ(r.iterations, 0u);
    CHECK_EQ(r.improved_count, 0u);
    CHECK_EQ(r.accepted_count, 0u);
    CHECK_EQ(r.initial_cost, r.final_cost);
}

TEST(test_optimization_single_iteration) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 1;
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK_EQ(r.iterations, 1u);
    CHECK(std::isfinite(r.final_cost));
}

TEST(test_optimization_large_grid) {
    SimilarityDatabase db;
    for (int i = 0; i < 20; ++i)
        db.add_tile("tile" + std::to_string(i) + ".png", lab(i * 5.0f, i * 2.0f - 20.0f, 10.0f - i));
    db.build_similarities();
    Grid grid(10, std::vector<std::optional<Path>>(10));
    for (int y = 0; y < 10; ++y)
        for (int x = 0; x < 10; ++x) grid[y][x] = "tile" + std::to_string((y * 10 + x) % 20) + ".png";
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 500;
    const float before = calc.calculate_total_cost(grid);
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK_EQ(r.iterations, 500u);
    CHECK_EQ(r.initial_cost, before);
    CHECK(r.best_cost <= r.initial_cost);
    // the returned running cost equals the recomputed cost of the final grid
    CHECK_NEAR(r.final_cost, calc.calculate_total_cost(grid), 1e-2);
}

TEST(test_greedy_optimization_edge_cases) {
    Grid empty;
    SimilarityDatabase db;
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    auto r = MosaicOptimizer(calc, OptimizationConfig{}).optimize_greedy(empty, 100);
    CHECK_EQ(r.initial_cost, 0.0f);
    CHECK_EQ(r.final_cost, 0.0f);
}

TEST(test_greedy_optimization_zero_iterations) {
    auto [grid, db] = create_test_grid();
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    auto r = MosaicOptimizer(calc, OptimizationConfig{}).optimize_greedy(grid, 0);
    CHECK_EQ(r.iterations, 0u);
    CHECK_EQ(r.initial_cost, r.final_cost);
}

HOST_TEST(test_optimization_result_edge_cases) {
    OptimizationResult zero;
    CHECK_EQ(zero.improvement_percentage(), 0.0f);
    OptimizationResult worse;
    worse.initial_cost = 50.0f;
    worse.final_cost = 75.0f;
    CHECK_EQ(worse.improvement_percentage(), -50.0f);
    OptimizationResult perfect;
    perfect.initial_cost = 100.0f;
    perfect.final_cost = 0.0f;
    CHECK_EQ(perfect.improvement_percentage(), 100.0f);
}

TEST(test_optimization_identical_tiles) {
    SimilarityDatabase db;
    for (int i = 0; i < 4; ++i) db.add_tile("same" + std::to_string(i) + ".png", lab(50, 0, 0));
    db.build_similarities();
    Grid grid(2, std::vector<std::optional<Path>>(2));
    grid[0][0] = "same0.png";
    grid[0][1] = "same1.png";
    grid[1][0] = "same2.png";
    grid[1][1] = "same3.png";
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 100;
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK_EQ(r.initial_cost, 4.0f);  // four adjacent pairs of distance 0
    CHECK_EQ(r.final_cost, r.initial_cost);  // every swap has delta 0
}

HOST_TEST(test_optimization_config_defaults) {
    OptimizationConfig c;
    CHECK_EQ(c.max_iterations, 1000u);
    CHECK_EQ(c.initial_temperature, 100.0f);
    CHECK_EQ(c.temperature_decay, 0.99995f);
    CHECK_EQ(c.report_interval, 100u);
}

TEST(test_optimization_numerical_stability) {
    SimilarityDatabase db;
    db.add_tile("a.png", lab(0, 0, 0));
    db.add_tile("b.png", lab(100, 127, 127));
    db.add_tile("c.png", lab(0.0001f, 0, 0));
    db.add_tile("d.png", lab(99.9999f, -128, -128));
    db.build_similarities();
    Grid grid(2, std::vector<std::optional<Path>>(2));
    grid[0][0] = "a.png";
    grid[0][1] = "b.png";
    grid[1][0] = "c.png";
    grid[1][1] = "d.png";
    AdjacencyPenaltyCalculator calc(db, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 100;
    auto r = MosaicOptimizer(calc, cfg).optimize(grid);
    CHECK(std::isfinite(r.initial_cost) && std::isfinite(r.final_cost) && std::isfinite(r.best_cost));
}

TEST(test_optimization_is_seeded) {
    auto [g1, db1] = create_test_grid();
    auto [g2, db2] = create_test_grid();
    AdjacencyPenaltyCalculator c1(db1, 1.0f), c2(db2, 1.0f);
    OptimizationConfig cfg;
    cfg.max_iterations = 200;
    cfg.seed = 42;
    auto r1 = MosaicOptimizer(c1, cfg).optimize(g1);
    auto r2 = MosaicOptimizer(c2, cfg).optimize(g2);
    CHECK(g1 == g2);
    CHECK_EQ(r1.final_cost, r2.final_cost);
    CHECK_EQ(r1.accepted_count, r2.accepted_count);
}

// ======================================================= color_adjustment.rs
HOST_TEST(test_default_adjustment) {
    ColorAdjustment a;
    CHECK_EQ(a.brightness, 0.0f);
    CHECK_EQ(a.contrast, 1.0f);
    CHECK_EQ(a.hue_shift, 0.0f);
    CHECK_EQ(a.saturation, 1.0f);
}

HOST_TEST(test_adjustment_clamping) {
    ColorAdjustment a = ColorAdjustment::create(2.0f, 3.0f, 200.0f, 3.0f);
    CHECK_EQ(a.brightness, 1.0f);
    CHECK_EQ(a.contrast, 2.0f);
    CHECK_EQ(a.hue_shift, 180.0f);
    CHECK_EQ(a.saturation, 2.0f);
}

TEST(test_pixel_adjustment_identity) {
    ColorAdjustment a;
    const Rgb px{100, 150, 200};
    const Rgb out = a.adjust_pixel(px);
    CHECK(std::abs((int)out.r - 100) <= 1);
    CHECK(std::abs((int)out.g - 150) <= 1);
    CHECK(std::abs((int)out.b - 200) <= 1);
}

TEST(test_brightness_adjustment) {
    ColorAdjustment a = ColorAdjustment::create(0.2f, 1.0f, 0.0f, 1.0f);
    const Rgb out = a.adjust_pixel(Rgb{100, 100, 100});
    CHECK(out.r > 100 && out.g > 100 && out.b > 100);
}

TEST(test_contrast_adjustment) {
    ColorAdjustment a = ColorAdjustment::create(0.0f, 1.5f, 0.0f, 1.0f);
    const Rgb dark = a.adjust_pixel(Rgb{64, 64, 64});
    const Rgb bright = a.adjust_pixel(Rgb{192, 192, 192});
    CHECK(dark.r < 64);
    CHECK(bright.r > 192);
}

TEST(test_optimal_adjustment_same_color) {
    ColorAdjustment a = calculate_optimal_adjustment(Rgb{128, 128, 128}, Rgb{128, 128, 128}, 1.0f);
    CHECK(std::fabs(a.brightness) < 0.01f);
    CHECK(std::fabs(a.hue_shift) < 0.01f);
    CHECK(std::fabs(a.saturation - 1.0f) < 0.01f);
}

TEST(test_pixel_adjustment_brightness_extremes) {
    const Rgb white = ColorAdjustment::create(1.0f, 1.0f, 0.0f, 1.0f).adjust_pixel(Rgb{128, 128, 128});
    CHECK(white == (Rgb{255, 255, 255}));
    const Rgb black = ColorAdjustment::create(-1.0f, 1.0f, 0.0f, 1.0f).adjust_pixel(Rgb{128, 128, 128});
    CHECK(black == (Rgb{0, 0, 0}));
}

TEST(test_pixel_adjustment_hue_shift) {
    // pure red shifted by +120 degrees becomes green
    const Rgb out = ColorAdjustment::create(0.0f, 1.0f, 120.0f, 1.0f).adjust_pixel(Rgb{255, 0, 0});
    CHECK(out.g > out.r && out.g > out.b);
}

TEST(test_pixel_adjustment_saturation) {
    const Rgb gray = ColorAdjustment::create(0.0f, 1.0f, 0.0f, 0.0f).adjust_pixel(Rgb{200, 100, 50});
    CHECK(std::abs((int)gray.r - (int)gray.g) <= 1 && std::abs((int)gray.g - (int)gray.b) <= 1);
}

TEST(test_apply_to_image) {
    RgbImage img(4, 3, Rgb{100, 100, 100});
    RgbImage out = ColorAdjustment::create(0.1f, 1.0f, 0.0f, 1.0f).apply_to_image(img);
    CHECK_EQ(out.width, 4u);
    CHECK_EQ(out.height, 3u);
    for (uint32_t y = 0; y < 3; ++y)
        for (uint32_t x = 0; x < 4; ++x) CHECK(out.get_pixel(x, y).r > 100);
}

TEST(test_optimal_adjustment_brightness_difference) {
    ColorAdjustment a = calculate_optimal_adjustment(Rgb{100, 100, 100}, Rgb{200, 200, 200}, 1.0f);
    CHECK(a.brightness > 0.0f);
    ColorAdjustment b = calculate_optimal_adjustment(Rgb{200, 200, 200}, Rgb{100, 100, 100}, 1.0f);
    CHECK(b.brightness < 0.0f);
}

TEST(test_optimal_adjustment_strength_scaling) {
    ColorAdjustment full = calculate_optimal_adjustment(Rgb{100, 100, 100}, Rgb{200, 200, 200}, 1.0f);
    ColorAdjustment half = calculate_optimal_adjustment(Rgb{100, 100, 100}, Rgb{200, 200, 200}, 0.5f);
    CHECK_NEAR(half.brightness, full.brightness * 0.5f, 1e-6);
}

TEST(test_optimal_adjustment_zero_strength) {
    ColorAdjustment a = calculate_optimal_adjustment(Rgb{50, 100, 150}, Rgb{200, 30, 90}, 0.0f);
    CHECK_EQ(a.brightness, 0.0f);
    CHECK_EQ(a.hue_shift, 0.0f);
    CHECK_EQ(a.saturation, 1.0f);
}

HOST_TEST(test_color_adjustment_new_negative_hue) {
    ColorAdjustment a = ColorAdjustment::create(0.0f, 1.0f, -200.0f, 1.0f);
    CHECK_EQ(a.hue_shift, -180.0f);
}

HOST_TEST(test_color_adjustment_new_extreme_values) {
    ColorAdjustment a = ColorAdjustment::create(-5.0f, -1.0f, 0.0f, -1.0f);
    CHECK_EQ(a.brightness, -1.0f);
    CHECK_EQ(a.contrast, 0.0f);
    CHECK_EQ(a.saturation, 0.0f);
}

TEST(test_extreme_pixel_values) {
    ColorAdjustment a = ColorAdjustment::create(0.5f, 2.0f, 90.0f, 2.0f);
    const Rgb w = a.adjust_pixel(Rgb{255, 255, 255});
    const Rgb b = a.adjust_pixel(Rgb{0, 0, 0});
    (void)w;
    (void)b;  // completes without overflow; u8 outputs are in range by type
}

// =================================================================== lib.rs
TEST(test_calculate_average_lab_single_color) {
    RgbImage img(10, 10, Rgb{255, 0, 0});
    Lab l = MosaicGeneratorImpl::calculate_average_lab(img);
    CHECK(std::fabs(l.l - 53.24f) < 1.0f);
    CHECK(std::fabs(l.a - 80.09f) < 1.0f);
    CHECK(std::fabs(l.b - 67.20f) < 1.0f);
}

TEST(test_calculate_average_lab_grayscale) {
    RgbImage img(10, 10, Rgb{128, 128, 128});
    Lab l = MosaicGeneratorImpl::calculate_average_lab(img);
    CHECK(std::fabs(l.l - 53.59f) < 1.0f);
    CHECK(std::fabs(l.a) < 1.0f);
    CHECK(std::fabs(l.b) < 1.0f);
}

HOST_TEST(test_aspect_ratio_match) {
    CHECK(MosaicGeneratorImpl::is_aspect_ratio_match(16.0f / 9.0f, 16.0f / 9.0f, 0.1f));
    CHECK(MosaicGeneratorImpl::is_aspect_ratio_match(1.8f, 16.0f / 9.0f, 0.1f));
    CHECK(!MosaicGeneratorImpl::is_aspect_ratio_match(2.0f, 16.0f / 9.0f, 0.1f));
    CHECK(!MosaicGeneratorImpl::is_aspect_ratio_match(9.0f / 16.0f, 16.0f / 9.0f, 0.1f));
}

HOST_TEST(test_usage_tracker) {
    UsageTracker t(2);
    const Path p = "test.png";
    CHECK(t.can_use_image(p));
    CHECK_EQ(t.get_usage_count(p), 0u);
    t.use_image(p);
    CHECK(t.can_use_image(p));
    CHECK_EQ(t.get_usage_count(p), 1u);
    t.use_image(p);
    CHECK_EQ(t.get_usage_count(p), 2u);
    CHECK(!t.can_use_image(p));
    CHECK(t.can_use_image("other.png"));
    t.reset();
    CHECK(t.can_use_image(p));
    CHECK_EQ(t.get_usage_count(p), 0u);
    UsageTracker zero(0);
    CHECK(!zero.can_use_image(p));
}

// nearest material for a red query is the red material (src/lib.rs:213-227),
// through the whole-job path: a 1x1 grid over a red target picks red
TEST(test_kdtree_nearest_neighbor_via_generate) {
    std::vector<RgbImage> mats = {RgbImage(8, 8, Rgb{0, 0, 255}), RgbImage(8, 8, Rgb{255, 0, 0}),
                                  RgbImage(8, 8, Rgb{0, 255, 0})};
    MaterialLibrary lib(mats);
    MosaicSettings s;
    s.grid_w = 1;
    s.grid_h = 1;
    s.color_adjustment_strength = 0.0f;
    s.enable_optimization = false;
    MosaicOutput out = lib.generate_mosaic(RgbImage(8, 8, Rgb{250, 5, 5}), s);
    CHECK_EQ(out.grid.size(), 1u);
    CHECK_EQ(out.grid[0], 1);
    CHECK(out.image.get_pixel(3, 3) == (Rgb{255, 0, 0}));
}

TEST(test_generate_mosaic_whole_job) {
    std::vector<RgbImage> mats;
    for (int i = 0; i < 12; ++i) mats.emplace_back(16, 16, Rgb{(uint8_t)(i * 20), (uint8_t)(255 - i * 20), 128});
    MaterialLibrary lib(mats);
    CHECK_EQ(lib.size(), 12);
    RgbImage target(96, 64);
    for (uint32_t y = 0; y < 64; ++y)
        for (uint32_t x = 0; x < 96; ++x) target.put_pixel(x, y, Rgb{(uint8_t)(x * 2), (uint8_t)(y * 3), 100});
    MosaicSettings s;
    s.grid_w = 6;
    s.grid_h = 4;
    s.max_usage_per_image = 3;
    s.optimization_iterations = 200;
    MosaicOutput out = lib.generate_mosaic(target, s);
    CHECK_EQ(out.result.place.placed, 24);
    CHECK_EQ(out.result.out_width, 96);
    CHECK_EQ(out.result.out_height, 64);
    for (int32_t m : out.grid) CHECK(m >= 0 && m < 12);
}

// ================================================= main.rs: material ingest
using Sizes = std::vector<std::optional<std::pair<uint32_t, uint32_t>>>;

Sizes squares(size_t n, uint32_t side = 100) {
    return Sizes(n, std::make_pair(side, side));
}

std::vector<size_t> iota_n(size_t n) {
    std::vector<size_t> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = i;
    return v;
}

// src/main.rs:1697-1727
HOST_TEST(test_load_tiles_file_extension_filtering) {
    for (const char* p : {"red.png", "red.jpg", "red.jpeg", "dir/RED.PNG", "x.WebP", "a.b.jpg"})
        CHECK(is_material_file(p));
    for (const char* p : {"text.txt", "data.dat", "png", ".png", "dir.png/file", "file.", "x.gif"})
        CHECK(!is_material_file(p));
}

// src/main.rs:1060-1078, :1093-1106, :1108-1121
HOST_TEST(test_load_tiles_selection_basic_limit_and_fallback) {
    MaterialSelection s = select_materials(squares(3), 1.0f, 0.1f, 10);
    CHECK(s.picked == iota_n(3) && !s.fallback);
    s = select_materials(squares(3), 1.0f, 0.1f, 2);
    CHECK(s.picked == iota_n(2));
    s = select_materials(squares(3), 10.0f, 0.1f, 10);
    CHECK(s.picked == iota_n(3) && s.fallback);
}

// src/main.rs:1730-1753, :1861-1886: undecodable entries are skipped
HOST_TEST(test_load_tiles_corrupted_image_handling) {
    Sizes sz = squares(2);
    sz[0] = std::nullopt;
    CHECK(select_materials(sz, 1.0f, 0.1f, 10).picked == std::vector<size_t>{1});
    CHECK(select_materials(sz, 7.0f, 0.1f, 10).picked == std::vector<size_t>{1});
}

// src/main.rs:1755-1777: the unfiltered sample is the first 2*max entries
HOST_TEST(test_load_tiles_aspect_ratio_fallback_sampling) {
    MaterialSelection s = select_materials(squares(10), 3.0f, 0.1f, 5);
    CHECK(s.fallback && s.picked == iota_n(5));
    Sizes sz = squares(10);
    sz[1] = sz[3] = std::nullopt;  // the sample (entries 0-5) holds 4 decodable ones
    s = select_materials(sz, 3.0f, 0.1f, 3);
    CHECK((s.picked == std::vector<size_t>{0, 2, 4}));
    s = select_materials(sz, 3.0f, 0.1f, 4);
    CHECK((s.picked == std::vector<size_t>{0, 2, 4, 5}));
}

// src/main.rs:1779-1796, :1825-1838, :2013-2022
HOST_TEST(test_load_tiles_max_materials_and_empty) {
    CHECK(select_materials(squares(20), 1.0f, 0.1f, 10).picked == iota_n(10));
    CHECK(select_materials(Sizes{}, 1.0f, 0.1f, 10).picked.empty());
    CHECK(select_materials(squares(3), 1.0f, 0.1f, 0).picked.empty());
    CHECK(select_materials(squares(3), 5.0f, 0.1f, 0).picked.empty());
}

// src/main.rs:1987-2011: the aspect test of process_tile; aspect = w/h in f32
HOST_TEST(test_load_tiles_aspect_tolerance_edges) {
    Sizes sz = {std::make_pair(100u, 100u), std::make_pair(200u, 100u), std::make_pair(177u, 100u)};
    CHECK(select_materials(sz, 1.0f, -0.1f, 10).fallback);  // nothing matches a negative tolerance
    CHECK(select_materials(sz, 1.0f, 1000.0f, 10).picked == iota_n(3));
    CHECK((select_materials(sz, 16.0f / 9.0f, 0.1f, 10).picked == std::vector<size_t>{2}));
    CHECK((select_materials(sz, 2.0f, 0.0f, 10).picked == std::vector<size_t>{1}));
}

// src/main.rs:808-827, :1798-1823, :2291-2305
TEST(test_load_tiles_lab_and_aspect) {
    std::vector<MaterialSource> e;
    e.push_back({"red.png", RgbImage(100, 100, Rgb{255, 0, 0})});
    e.push_back({"broken.png", std::nullopt});
    for (int i = 0; i < 5; ++i)
        e.push_back({"img_" + std::to_string(i) + ".png",
                     RgbImage(100, 100, Rgb{(uint8_t)(i * 50 + 50), 0, 0})});
    e.push_back({"tiny.png", RgbImage(1, 1, Rgb{255, 255, 255})});
    e.push_back({"wide.png", RgbImage(300, 100, Rgb{0, 0, 255})});
    std::vector<Tile> t = load_tiles(e, 1.0f, 0.1f, 10);
    CHECK_EQ(t.size(), 7u);  // broken and wide are out
    CHECK_EQ(t[0].path, "red.png");
    CHECK(std::fabs(t[0].lab_color.l - 53.0f) < 5.0f);
    CHECK(std::fabs(t[0].lab_color.a - 80.0f) < 5.0f);
    CHECK(std::fabs(t[0].lab_color.b - 67.0f) < 5.0f);
    for (const Tile& x : t) {
        CHECK_EQ(x.aspect_ratio, 1.0f);
        CHECK(x.lab_color.l > 0.0f && x.lab_color.l <= 100.0f);
        CHECK_EQ(e[x.source_index].path, x.path);
    }
    CHECK_EQ(t[6].path, "tiny.png");
}

// the batched ingest equals calculate_average_lab image by image, for
// images of several sizes in mixed order
TEST(test_load_tiles_mixed_sizes_match_single_image_lab) {
    std::vector<MaterialSource> e;
    const uint32_t dims[][2] = {{64, 64}, {3, 7}, {100, 37}, {64, 64}, {1, 1}, {5, 300}, {3, 7}};
    uint32_t seed = 7;
    for (auto& d : dims) {
        RgbImage im(d[0], d[1]);
        for (uint8_t& b : im.data) b = (uint8_t)((seed = seed * 1103515245u + 12345u) >> 24);
        e.push_back({"m" + std::to_string(e.size()) + ".png", im});
    }
    std::vector<Tile> t = load_tiles(e, 1.0f, 1000.0f, 100);
    CHECK_EQ(t.size(), e.size());
    for (size_t i = 0; i < t.size(); ++i) {
        const Lab one = MosaicGeneratorImpl::calculate_average_lab(*e[i].image);
        CHECK(std::memcmp(&one, &t[i].lab_color, sizeof(Lab)) == 0);
    }
}

// src/main.rs:1561-1585, :1587-1624, :1656-1695
TEST(test_new_with_similarity_database_rebuild_load_and_missing_tiles) {
    std::vector<MaterialSource> e;
    e.push_back({"red.png", RgbImage(100, 100, Rgb{255, 0, 0})});
    e.push_back({"green.png", RgbImage(100, 100, Rgb{0, 255, 0})});
    e.push_back({"blue.png", RgbImage(100, 100, Rgb{0, 0, 255})});
    const std::string path = temp_path("prepare");
    std::remove(path.c_str());
    std::vector<Tile> t = load_tiles(e, 1.0f, 0.1f, 10);
    SimilarityDatabase db = prepare_similarity_db(t, path, true);
    CHECK_EQ(db.size(), 3u);
    CHECK(std::ifstream(path).good());
    // an existing file is loaded, not rebuilt
    SimilarityDatabase db2 = prepare_similarity_db(t, path, false);
    CHECK_EQ(db2.size(), 3u);
    CHECK(db2.get_similarity("red.png", "blue.png") == db.get_similarity("red.png", "blue.png"));
    // a new tile is added to the loaded database
    e.push_back({"yellow.png", RgbImage(100, 100, Rgb{255, 255, 0})});
    t = load_tiles(e, 1.0f, 0.1f, 10);
    CHECK_EQ(t.size(), 4u);
    SimilarityDatabase db3 = prepare_similarity_db(t, path, false);
    CHECK_EQ(db3.size(), 4u);
    CHECK(db3.get_similarity("yellow.png", "red.png").has_value());
    std::remove(path.c_str());
}

// materials of different sizes in one library: each is resized from its own
// size, so a solid material composites to its colour whatever its size
TEST(test_generate_mosaic_mixed_size_materials) {
    std::vector<RgbImage> mats = {RgbImage(40, 40, Rgb{255, 0, 0}), RgbImage(7, 13, Rgb{0, 255, 0}),
                                  RgbImage(64, 48, Rgb{0, 0, 255}), RgbImage(7, 13, Rgb{250, 250, 0})};
    MaterialLibrary lib(mats);
    RgbImage target(64, 16);
    for (uint32_t y = 0; y < 16; ++y)
        for (uint32_t x = 0; x < 64; ++x) {
            const Rgb c[4] = {{255, 0, 0}, {0, 255, 0}, {0, 0, 255}, {250, 250, 0}};
            target.put_pixel(x, y, c[x / 16]);
        }
    MosaicSettings s;
    s.grid_w = 4;
    s.grid_h = 1;
    s.color_adjustment_strength = 0.0f;
    s.enable_optimization = false;
    MosaicOutput out = lib.generate_mosaic(target, s);
    CHECK((out.grid == std::vector<int32_t>{0, 1, 2, 3}));
    for (uint32_t x = 0; x < 64; ++x) CHECK(out.image.get_pixel(x, 5) == target.get_pixel(x, 5));
}

int main(int argc, char** argv) {
    bool host_only = false, list = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--host")) host_only = true;
        if (!std::strcmp(argv[i], "--list")) list = true;
    }
    int failed = 0, ran = 0;
    for (const TestCase& t : registry()) {
        if (list) {
            std::printf("%s%s\n", t.name, t.host_only ? " [host]" : "");
            continue;
        }
        if (host_only && !t.host_only) continue;
        ++ran;
        try {
            t.fn();
            std::printf("PASS %s\n", t.name);
        } catch (const Failure& f) {
            ++failed;
            std::printf("FAIL %s: %s\n", t.name, f.msg.c_str());
        } catch (const std::exception& e) {
            ++failed;
            std::printf("FAIL %s: exception: %s\n", t.name, e.what());
        }
        std::fflush(stdout);
    }
    if (!list) std::printf("%d tests, %d failed\n", ran, failed);
    return failed ? 1 : 0;
}
