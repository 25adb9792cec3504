// mosaic.cpp -- C++ host API over the C ABI (see mosaic.hpp).
#include "mosaic.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <mutex>
#include <sstream>

namespace mosaic {

namespace {

void check(int rc) {
    if (rc != MH_OK) throw Error(rc, std::string("mosaic_hip: ") + mh_last_error());
}

std::shared_ptr<Context> or_default(std::shared_ptr<Context> ctx) {
    return ctx ? ctx : Context::default_context();
}

// The device kernels use every pairwise distance; the reference looks
// distances up in its built table (src/similarity.rs:98-102), which is empty
// until build_similarities.  Refuse that state instead of diverging.
void require_built(const SimilarityDatabase& db) {
    const size_t n = db.size();
    if (n >= 2 && db.similarities().size() != n * (n - 1) / 2)
        throw Error(MH_ERR_INVALID_ARGUMENT, "similarities not built: call build_similarities() first");
}

// Index grid (row-major, -1 = empty or unknown path) of a path grid.
std::vector<int32_t> to_index_grid(const SimilarityDatabase& db, const Grid& grid, size_t* w,
                                   size_t* h) {
    *h = grid.size();
    *w = *h ? grid[0].size() : 0;
    std::vector<int32_t> g(*w * *h, MH_EMPTY);
    for (size_t y = 0; y < *h; ++y) {
        if (grid[y].size() != *w) throw Error(MH_ERR_INVALID_ARGUMENT, "grid rows differ in length");
        for (size_t x = 0; x < *w; ++x) {
            if (!grid[y][x]) continue;
            const auto idx = db.index_of(*grid[y][x]);
            if (idx) g[y * *w + x] = *idx;
        }
    }
    return g;
}

// ------------------------------------------------------------------ JSON
// The reference database is serde_json pretty output of
//   { path_to_index: {path: index}, index_to_path: {"index": path},
//     lab_colors: [{l, a, b}], similarities: [f32] }.
struct JValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    std::string text;  // number token or string contents
    std::vector<JValue> items;
    std::vector<std::pair<std::string, JValue>> fields;
    const JValue* get(const std::string& k) const {
        for (const auto& f : fields)
            if (f.first == k) return &f.second;
        return nullptr;
    }
};

struct JParser {
    const std::string& s;
    size_t i = 0;
    [[noreturn]] void fail(const char* what) {
        throw Error(MH_ERR_INVALID_ARGUMENT,
                    std::string("similarity database JSON: ") + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
    }
    void expect(char c) {
        ws();
        if (i >= s.size() || s[i] != c) fail("unexpected character");
        ++i;
    }
    static void put_utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (i + 4 > s.size()) fail("short \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        expect('"');
        std::string out;
        while (true) {
            if (i >= s.size()) fail("unterminated string");
            const char c = s[i++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("unterminated escape");
            const char e = s[i++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && i + 1 < s.size() && s[i] == '\\' && s[i + 1] == 'u') {
                        i += 2;
                        const uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    JValue value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        JValue v;
        const char c = s[i];
        if (c == '{') {
            v.kind = JValue::Object;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return v;
            }
            while (true) {
                std::string k = str();
                expect(':');
                v.fields.emplace_back(std::move(k), value());
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                expect('}');
                return v;
            }
        }
        if (c == '[') {
            v.kind = JValue::Array;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return v;
            }
            while (true) {
                v.items.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                expect(']');
                return v;
            }
        }
        if (c == '"') {
            v.kind = JValue::String;
            v.text = str();
            return v;
        }
        if (s.compare(i, 4, "true") == 0) {
            i += 4;
            v.kind = JValue::Bool;
            v.b = true;
            return v;
        }
        if (s.compare(i, 5, "false") == 0) {
            i += 5;
            v.kind = JValue::Bool;
            return v;
        }
        if (s.compare(i, 4, "null") == 0) {
            i += 4;
            return v;
        }
        const size_t st = i;
        while (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '-' || s[i] == '+' ||
                                s[i] == '.' || s[i] == 'e' || s[i] == 'E'))
            ++i;
        if (i == st) fail("unexpected character");
        v.kind = JValue::Number;
        v.text = s.substr(st, i - st);
        return v;
    }
};

float num_f32(const JValue& v) {
    if (v.kind != JValue::Number) throw Error(MH_ERR_INVALID_ARGUMENT, "expected a number");
    return std::strtof(v.text.c_str(), nullptr);  // correctly rounded to f32
}

size_t num_index(const std::string& t) {
    char* end = nullptr;
    const unsigned long long v = std::strtoull(t.c_str(), &end, 10);
    if (t.empty() || *end != 0 || t[0] == '-') throw Error(MH_ERR_INVALID_ARGUMENT, "expected an index");
    return (size_t)v;
}

void json_string(std::string& out, const std::string& s) {
    out += '"';
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            default:
                if (c < 0x20) {
                    char b[8];
                    std::snprintf(b, sizeof(b), "\\u%04x", c);
                    out += b;
                } else {
                    out += (char)c;
                }
        }
    }
    out += '"';
}

// f32 as serde_json writes it: the shortest round-trip digits, laid out like
// the ryu crate's pretty printer (decimal for 10^-6 <= |v| < 10^13, else
// d.ddde[-]x; integral values end in ".0").
void json_f32(std::string& out, float f) {
    if (!std::isfinite(f)) {
        out += "null";
        return;
    }
    if (f == 0.0f) {
        out += std::signbit(f) ? "-0.0" : "0.0";
        return;
    }
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof(buf), f, std::chars_format::scientific);
    std::string sci(buf, r.ptr);
    std::string sign;
    if (sci[0] == '-') {
        sign = "-";
        sci.erase(0, 1);
    }
    const size_t epos = sci.find('e');
    std::string mant = sci.substr(0, epos);
    const int e10 = std::atoi(sci.c_str() + epos + 1);
    std::string digits;
    for (char c : mant)
        if (c != '.') digits += c;
    const int length = (int)digits.size();
    const int k = e10 - (length - 1);  // value = digits * 10^k
    const int kk = length + k;
    std::string s;
    if (k >= 0 && kk <= 13) {
        s = digits + std::string(k, '0') + ".0";
    } else if (kk > 0 && kk <= 13) {
        s = digits.substr(0, kk) + "." + digits.substr(kk);
    } else if (kk > -6 && kk <= 0) {
        s = "0." + std::string(-kk, '0') + digits;
    } else if (length == 1) {
        s = digits + "e" + std::to_string(kk - 1);
    } else {
        s = digits.substr(0, 1) + "." + digits.substr(1) + "e" + std::to_string(kk - 1);
    }
    out += sign + s;
}

}  // namespace

// --------------------------------------------------------------- context
Context::Context(int32_t device) { check(mh_context_create(device, &h_)); }

Context::~Context() {
    if (h_) mh_context_destroy(h_);
}

std::shared_ptr<Context> Context::default_context() {
    static std::mutex mu;
    static std::shared_ptr<Context> ctx;
    std::lock_guard<std::mutex> g(mu);
    if (!ctx) ctx = std::make_shared<Context>(0);
    return ctx;
}

RgbImage::RgbImage(uint32_t w, uint32_t h, Rgb fill) : width(w), height(h), data((size_t)w * h * 3) {
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        data[3 * i] = fill.r;
        data[3 * i + 1] = fill.g;
        data[3 * i + 2] = fill.b;
    }
}

Rgb RgbImage::get_pixel(uint32_t x, uint32_t y) const {
    const size_t o = ((size_t)y * width + x) * 3;
    return Rgb{data.at(o), data.at(o + 1), data.at(o + 2)};
}

void RgbImage::put_pixel(uint32_t x, uint32_t y, Rgb p) {
    const size_t o = ((size_t)y * width + x) * 3;
    data.at(o) = p.r;
    data.at(o + 1) = p.g;
    data.at(o + 2) = p.b;
}

// ------------------------------------------------------------ similarity
SimilarityDatabase::SimilarityDatabase() = default;

SimilarityDatabase::SimilarityDatabase(std::shared_ptr<Context> ctx) : ctx_(std::move(ctx)) {}

SimilarityDatabase::~SimilarityDatabase() { drop_library(); }

SimilarityDatabase::SimilarityDatabase(const SimilarityDatabase& o)
    : ctx_(o.ctx_),
      path_to_index_(o.path_to_index_),
      index_to_path_(o.index_to_path_),
      lab_colors_(o.lab_colors_),
      similarities_(o.similarities_) {}

SimilarityDatabase& SimilarityDatabase::operator=(const SimilarityDatabase& o) {
    if (this != &o) {
        drop_library();
        ctx_ = o.ctx_;
        path_to_index_ = o.path_to_index_;
        index_to_path_ = o.index_to_path_;
        lab_colors_ = o.lab_colors_;
        similarities_ = o.similarities_;
    }
    return *this;
}

SimilarityDatabase::SimilarityDatabase(SimilarityDatabase&& o) noexcept
    : ctx_(std::move(o.ctx_)),
      path_to_index_(std::move(o.path_to_index_)),
      index_to_path_(std::move(o.index_to_path_)),
      lab_colors_(std::move(o.lab_colors_)),
      similarities_(std::move(o.similarities_)),
      lib_(o.lib_),
      lib_n_(o.lib_n_) {
    o.lib_ = nullptr;
}

SimilarityDatabase& SimilarityDatabase::operator=(SimilarityDatabase&& o) noexcept {
    if (this != &o) {
        drop_library();
        ctx_ = std::move(o.ctx_);
        path_to_index_ = std::move(o.path_to_index_);
        index_to_path_ = std::move(o.index_to_path_);
        lab_colors_ = std::move(o.lab_colors_);
        similarities_ = std::move(o.similarities_);
        lib_ = o.lib_;
        lib_n_ = o.lib_n_;
        o.lib_ = nullptr;
    }
    return *this;
}

void SimilarityDatabase::drop_library() {
    if (lib_) mh_library_destroy(lib_);
    lib_ = nullptr;
    lib_n_ = 0;
}

// src/similarity.rs:61-66: the new tile takes the next index; a repeated path
// is re-pointed at its newest index.
void SimilarityDatabase::add_tile(const Path& path, const Lab& lab_color) {
    const size_t idx = lab_colors_.size();
    path_to_index_[path] = idx;
    index_to_path_[idx] = path;
    lab_colors_.push_back(lab_color);
}

mh_library* SimilarityDatabase::device_library() const {
    if (lib_ && lib_n_ == lab_colors_.size()) return lib_;
    if (lib_) mh_library_destroy(lib_);
    lib_ = nullptr;
    auto* self = const_cast<SimilarityDatabase*>(this);
    self->ctx_ = or_default(ctx_);
    std::vector<float> lab(lab_colors_.size() * 3);
    for (size_t i = 0; i < lab_colors_.size(); ++i) {
        lab[3 * i] = lab_colors_[i].l;
        lab[3 * i + 1] = lab_colors_[i].a;
        lab[3 * i + 2] = lab_colors_[i].b;
    }
    check(mh_library_from_lab(ctx_->get(), lab.data(), (int32_t)lab_colors_.size(), &lib_));
    lib_n_ = lab_colors_.size();
    return lib_;
}

void SimilarityDatabase::build_similarities() {
    const size_t n = lab_colors_.size();
    similarities_.assign(n > 1 ? n * (n - 1) / 2 : 0, 0.0f);
    if (n < 2) return;
    mh_library* lib = device_library();
    check(mh_library_similarities(lib, similarities_.data(), 1));
}

std::optional<int32_t> SimilarityDatabase::index_of(const Path& path) const {
    auto it = path_to_index_.find(path);
    if (it == path_to_index_.end()) return std::nullopt;
    return (int32_t)it->second;
}

const Path* SimilarityDatabase::path_of(int32_t index) const {
    auto it = index_to_path_.find((size_t)index);
    return it == index_to_path_.end() ? nullptr : &it->second;
}

// src/similarity.rs:85-103
std::optional<float> SimilarityDatabase::get_similarity(const Path& path1, const Path& path2) const {
    auto a = path_to_index_.find(path1);
    auto b = path_to_index_.find(path2);
    if (a == path_to_index_.end() || b == path_to_index_.end()) return std::nullopt;
    if (a->second == b->second) return 0.0f;
    const size_t i = std::min(a->second, b->second), j = std::max(a->second, b->second);
    const size_t n = lab_colors_.size();
    const size_t pos = i * n - i * (i + 1) / 2 + j - i - 1;
    if (pos >= similarities_.size()) return std::nullopt;
    return similarities_[pos];
}

std::optional<Lab> SimilarityDatabase::get_lab_color(const Path& path) const {
    auto it = path_to_index_.find(path);
    if (it == path_to_index_.end() || it->second >= lab_colors_.size()) return std::nullopt;
    return lab_colors_[it->second];
}

void SimilarityDatabase::save_to_file(const Path& path) const {
    std::string o = "{\n  \"path_to_index\": {";
    // index order (the reference's HashMap order is arbitrary)
    std::vector<std::pair<size_t, Path>> p2i;
    for (const auto& kv : path_to_index_) p2i.emplace_back(kv.second, kv.first);
    std::sort(p2i.begin(), p2i.end());
    for (size_t k = 0; k < p2i.size(); ++k) {
        o += k ? ",\n    " : "\n    ";
        json_string(o, p2i[k].second);
        o += ": " + std::to_string(p2i[k].first);
    }
    o += p2i.empty() ? "}" : "\n  }";
    o += ",\n  \"index_to_path\": {";
    size_t k = 0;
    for (const auto& kv : index_to_path_) {
        o += k++ ? ",\n    " : "\n    ";
        json_string(o, std::to_string(kv.first));
        o += ": ";
        json_string(o, kv.second);
    }
    o += index_to_path_.empty() ? "}" : "\n  }";
    o += ",\n  \"lab_colors\": [";
    for (size_t i = 0; i < lab_colors_.size(); ++i) {
        o += i ? ",\n    {\n      \"l\": " : "\n    {\n      \"l\": ";
        json_f32(o, lab_colors_[i].l);
        o += ",\n      \"a\": ";
        json_f32(o, lab_colors_[i].a);
        o += ",\n      \"b\": ";
        json_f32(o, lab_colors_[i].b);
        o += "\n    }";
    }
    o += lab_colors_.empty() ? "]" : "\n  ]";
    o += ",\n  \"similarities\": [";
    for (size_t i = 0; i < similarities_.size(); ++i) {
        o += i ? ",\n    " : "\n    ";
        json_f32(o, similarities_[i]);
    }
    o += similarities_.empty() ? "]" : "\n  ]";
    o += "\n}";
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error(MH_ERR_INVALID_ARGUMENT, "cannot open " + path + " for writing");
    f << o;
    if (!f) throw Error(MH_ERR_INVALID_ARGUMENT, "cannot write " + path);
}

SimilarityDatabase SimilarityDatabase::load_from_file(const Path& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(MH_ERR_NOT_FOUND, "cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JParser p{text};
    const JValue root = p.value();
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    if (root.kind != JValue::Object) p.fail("expected an object");
    const JValue* p2i = root.get("path_to_index");
    const JValue* i2p = root.get("index_to_path");
    const JValue* labs = root.get("lab_colors");
    const JValue* sims = root.get("similarities");
    if (!p2i || !i2p || !labs || !sims || p2i->kind != JValue::Object ||
        i2p->kind != JValue::Object || labs->kind != JValue::Array || sims->kind != JValue::Array)
        p.fail("missing or mistyped field");
    SimilarityDatabase db;
    for (const auto& kv : p2i->fields) db.path_to_index_[kv.first] = num_index(kv.second.text);
    for (const auto& kv : i2p->fields) {
        if (kv.second.kind != JValue::String) p.fail("index_to_path value is not a string");
        db.index_to_path_[num_index(kv.first)] = kv.second.text;
    }
    for (const auto& l : labs->items) {
        const JValue *jl = l.get("l"), *ja = l.get("a"), *jb = l.get("b");
        if (!jl || !ja || !jb) p.fail("lab colour without l, a, b");
        db.lab_colors_.push_back(Lab{num_f32(*jl), num_f32(*ja), num_f32(*jb)});
    }
    for (const auto& s : sims->items) db.similarities_.push_back(num_f32(s));
    return db;
}

SimilarityDatabase SimilarityDatabase::load_or_new(const Path& path) {
    try {
        return load_from_file(path);
    } catch (const Error&) {
        return SimilarityDatabase();
    }
}

// ------------------------------------------------------------- adjacency
std::vector<GridPosition> GridPosition::get_adjacent_positions(size_t grid_width,
                                                               size_t grid_height) const {
    std::vector<GridPosition> adj;
    if (y > 0) adj.emplace_back(x, y - 1);
    if (grid_height > 0 && y < grid_height - 1) adj.emplace_back(x, y + 1);
    if (x > 0) adj.emplace_back(x - 1, y);
    if (grid_width > 0 && x < grid_width - 1) adj.emplace_back(x + 1, y);
    return adj;
}

AdjacencyPenaltyCalculator::AdjacencyPenaltyCalculator(const SimilarityDatabase& similarity_db,
                                                       float penalty_weight)
    : db_(&similarity_db), penalty_weight_(penalty_weight) {}

float AdjacencyPenaltyCalculator::calculate_penalty(const Path& candidate_path, GridPosition position,
                                                    const Grid& grid, size_t grid_width,
                                                    size_t grid_height) const {
    const auto m = db_->index_of(candidate_path);
    if (!m || db_->size() == 0) return 0.0f * penalty_weight_;  // every lookup misses
    require_built(*db_);
    size_t w, h;
    std::vector<int32_t> g = to_index_grid(*db_, grid, &w, &h);
    // neighbours are taken from a grid_width x grid_height view (src/adjacency.rs:70)
    std::vector<int32_t> view(grid_width * grid_height, MH_EMPTY);
    for (size_t y = 0; y < grid_height && y < h; ++y)
        for (size_t x = 0; x < grid_width && x < w; ++x) view[y * grid_width + x] = g[y * w + x];
    if (grid_width == 0 || grid_height == 0) return 0.0f;
    const int32_t q[3] = {*m, (int32_t)position.x, (int32_t)position.y};
    float out = 0.0f;
    check(mh_penalties(db_->context()->get(), db_->device_library(), view.data(),
                       (int32_t)grid_width, (int32_t)grid_height, penalty_weight_, q, 1, &out));
    return out;
}

float AdjacencyPenaltyCalculator::calculate_total_cost(const Grid& grid) const {
    size_t w, h;
    std::vector<int32_t> g = to_index_grid(*db_, grid, &w, &h);
    if (h == 0 || w == 0) return 0.0f;
    require_built(*db_);
    mh_library* lib = db_->device_library();
    float out = 0.0f;
    check(mh_total_cost(db_->context()->get(), lib, g.data(), (int32_t)w, (int32_t)h, &out));
    return out;
}

float AdjacencyPenaltyCalculator::calculate_swap_delta(const Grid& grid, GridPosition pos1,
                                                       GridPosition pos2) const {
    size_t w, h;
    std::vector<int32_t> g = to_index_grid(*db_, grid, &w, &h);
    if (h == 0 || w == 0) return 0.0f;
    require_built(*db_);
    mh_library* lib = db_->device_library();
    const int32_t pr[4] = {(int32_t)pos1.x, (int32_t)pos1.y, (int32_t)pos2.x, (int32_t)pos2.y};
    float out = 0.0f;
    check(mh_swap_deltas(db_->context()->get(), lib, g.data(), (int32_t)w, (int32_t)h, pr, 1, &out));
    return out;
}

// ------------------------------------------------------------- optimizer
float OptimizationResult::improvement_percentage() const {
    if (initial_cost > 0.0f) return ((initial_cost - final_cost) / initial_cost) * 100.0f;
    return 0.0f;
}

MosaicOptimizer::MosaicOptimizer(const AdjacencyPenaltyCalculator& calculator,
                                 OptimizationConfig config)
    : calc_(&calculator), config_(config) {}

OptimizationResult MosaicOptimizer::run(Grid& grid, size_t iterations, bool greedy) const {
    const SimilarityDatabase& db = calc_->similarity_db();
    size_t w, h;
    std::vector<int32_t> g = to_index_grid(db, grid, &w, &h);
    if (h == 0 || w == 0) return OptimizationResult{};
    require_built(db);
    mh_library* lib = db.device_library();
    mh_anneal_params p{};
    p.max_iterations = (int64_t)iterations;
    p.initial_temperature = config_.initial_temperature;
    p.temperature_decay = config_.temperature_decay;
    p.seed = config_.seed;
    p.greedy = greedy ? 1 : 0;
    mh_anneal_result r{};
    check(mh_anneal(db.context()->get(), lib, g.data(), (int32_t)w, (int32_t)h, &p, &r, nullptr,
                    nullptr));
    // write the permuted grid back (only placed, known cells move)
    for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x) {
            const int32_t m = g[y * w + x];
            if (m >= 0) grid[y][x] = *db.path_of(m);
        }
    OptimizationResult out;
    out.initial_cost = r.initial_cost;
    out.final_cost = r.final_cost;
    out.best_cost = r.best_cost;
    out.improved_count = (size_t)r.improved_count;
    out.accepted_count = (size_t)r.accepted_count;
    out.iterations = (size_t)r.iterations;
    return out;
}

OptimizationResult MosaicOptimizer::optimize(Grid& grid) const {
    return run(grid, config_.max_iterations, false);
}

OptimizationResult MosaicOptimizer::optimize_greedy(Grid& grid, size_t max_iterations) const {
    return run(grid, max_iterations, true);
}

// ------------------------------------------------------ colour adjustment
ColorAdjustment ColorAdjustment::create(float brightness, float contrast, float hue_shift,
                                        float saturation) {
    auto clamp = [](float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); };
    ColorAdjustment a;
    a.brightness = clamp(brightness, -1.0f, 1.0f);
    a.contrast = clamp(contrast, 0.0f, 2.0f);
    a.hue_shift = clamp(hue_shift, -180.0f, 180.0f);
    a.saturation = clamp(saturation, 0.0f, 2.0f);
    return a;
}

RgbImage ColorAdjustment::apply_to_image(const RgbImage& img) const {
    RgbImage out(img.width, img.height);
    const float adj[4] = {brightness, contrast, hue_shift, saturation};
    const int64_t npx = (int64_t)img.width * img.height;
    if (npx > 0)
        check(mh_adjust_pixels(Context::default_context()->get(), adj, img.data.data(), npx,
                               out.data.data()));
    return out;
}

Rgb ColorAdjustment::adjust_pixel(Rgb pixel) const {
    RgbImage one(1, 1, pixel);
    return apply_to_image(one).get_pixel(0, 0);
}

ColorAdjustment calculate_optimal_adjustment(Rgb tile_avg_rgb, Rgb target_avg_rgb,
                                             float adjustment_strength) {
    const uint8_t t[3] = {tile_avg_rgb.r, tile_avg_rgb.g, tile_avg_rgb.b};
    const uint8_t g[3] = {target_avg_rgb.r, target_avg_rgb.g, target_avg_rgb.b};
    float out[4];
    check(mh_optimal_adjustments(Context::default_context()->get(), t, g, 1, adjustment_strength,
                                 out));
    ColorAdjustment a;
    a.brightness = out[0];
    a.contrast = out[1];
    a.hue_shift = out[2];
    a.saturation = out[3];
    return a;
}

// ------------------------------------------------------------------ lib
bool UsageTracker::can_use_image(const Path& path) const { return get_usage_count(path) < max_; }

void UsageTracker::use_image(const Path& path) { counts_[path] = get_usage_count(path) + 1; }

size_t UsageTracker::get_usage_count(const Path& path) const {
    auto it = counts_.find(path);
    return it == counts_.end() ? 0 : it->second;
}

Lab MosaicGeneratorImpl::calculate_average_lab(const RgbImage& img) {
    float out[3];
    check(mh_average_lab(Context::default_context()->get(), img.data.data(), (int32_t)img.width,
                         (int32_t)img.height, out));
    return Lab{out[0], out[1], out[2]};
}

bool MosaicGeneratorImpl::is_aspect_ratio_match(float img_aspect, float target_aspect,
                                                float tolerance) {
    return std::fabs(img_aspect - target_aspect) <= tolerance;
}

// ------------------------------------------------------- material ingest
namespace {

// The images as one buffer with per-image offsets and sizes (the layout of
// mh_average_lab_many / mh_library_from_images).
struct PackedImages {
    std::vector<uint8_t> bytes;
    std::vector<int64_t> offsets;
    std::vector<int32_t> widths, heights;
};

PackedImages pack_images(const std::vector<const RgbImage*>& images) {
    PackedImages p;
    size_t total = 0;
    for (const RgbImage* im : images) {
        if (im->data.size() != (size_t)im->width * im->height * 3)
            throw Error(MH_ERR_INVALID_ARGUMENT, "image data does not match its size");
        total += im->data.size();
    }
    p.bytes.reserve(total);
    for (const RgbImage* im : images) {
        p.offsets.push_back((int64_t)p.bytes.size());
        p.widths.push_back((int32_t)im->width);
        p.heights.push_back((int32_t)im->height);
        p.bytes.insert(p.bytes.end(), im->data.begin(), im->data.end());
    }
    return p;
}

}  // namespace

bool is_material_file(const Path& path) {
    const size_t slash = path.find_last_of('/');
    const std::string name = slash == Path::npos ? path : path.substr(slash + 1);
    const size_t dot = name.find_last_of('.');
    // Path::extension is None for a name that only starts with a dot (".png")
    if (dot == std::string::npos || dot == 0) return false;
    std::string ext = name.substr(dot + 1);
    for (char& c : ext) c = (char)std::tolower((unsigned char)c);
    return ext == "png" || ext == "jpg" || ext == "jpeg" || ext == "webp";
}

MaterialSelection select_materials(
    const std::vector<std::optional<std::pair<uint32_t, uint32_t>>>& sizes, float target_aspect,
    float aspect_tolerance, size_t max_materials) {
    MaterialSelection sel;
    for (size_t i = 0; i < sizes.size(); ++i) {  // process_tile, :263-285
        if (!sizes[i]) continue;
        const float aspect = (float)sizes[i]->first / (float)sizes[i]->second;
        if (MosaicGeneratorImpl::is_aspect_ratio_match(aspect, target_aspect, aspect_tolerance))
            sel.picked.push_back(i);
    }
    if (sel.picked.empty()) {  // :214-249: unfiltered sample of the first 2*max entries
        sel.fallback = true;
        const size_t take = std::min(sizes.size(), max_materials * 2);
        for (size_t i = 0; i < take; ++i)
            if (sizes[i]) sel.picked.push_back(i);
    }
    if (sel.picked.size() > max_materials) sel.picked.resize(max_materials);  // :252-254
    return sel;
}

std::vector<Tile> load_tiles(const std::vector<MaterialSource>& entries, float target_aspect,
                             float aspect_tolerance, size_t max_materials,
                             std::shared_ptr<Context> ctx) {
    std::vector<std::optional<std::pair<uint32_t, uint32_t>>> sizes(entries.size());
    for (size_t i = 0; i < entries.size(); ++i)
        if (entries[i].image && entries[i].image->width > 0 && entries[i].image->height > 0)
            sizes[i] = std::make_pair(entries[i].image->width, entries[i].image->height);
    const MaterialSelection sel = select_materials(sizes, target_aspect, aspect_tolerance,
                                                   max_materials);
    std::vector<Tile> tiles;
    if (sel.picked.empty()) return tiles;
    std::vector<const RgbImage*> images;
    for (size_t i : sel.picked) images.push_back(&*entries[i].image);
    const PackedImages p = pack_images(images);
    std::vector<float> lab(images.size() * 3);
    check(mh_average_lab_many(or_default(std::move(ctx))->get(), p.bytes.data(), p.offsets.data(),
                              p.widths.data(), p.heights.data(), (int32_t)images.size(),
                              lab.data()));
    for (size_t k = 0; k < sel.picked.size(); ++k) {
        const RgbImage& im = *images[k];
        Tile t;
        t.path = entries[sel.picked[k]].path;
        t.lab_color = Lab{lab[3 * k], lab[3 * k + 1], lab[3 * k + 2]};
        t.aspect_ratio = (float)im.width / (float)im.height;
        t.source_index = sel.picked[k];
        tiles.push_back(std::move(t));
    }
    return tiles;
}

SimilarityDatabase prepare_similarity_db(const std::vector<Tile>& tiles, const Path& db_path,
                                         bool rebuild_similarity) {
    SimilarityDatabase db;
    if (rebuild_similarity || !std::filesystem::exists(db_path)) {
        for (const Tile& t : tiles) db.add_tile(t.path, t.lab_color);
        db.build_similarities();
        try {
            db.save_to_file(db_path);
        } catch (const std::exception& e) {
            std::cerr << "Warning: Failed to save similarity database: " << e.what() << "\n";
        }
    } else {
        db = SimilarityDatabase::load_or_new(db_path);
    }
    for (const Tile& t : tiles)
        if (!db.get_lab_color(t.path)) db.add_tile(t.path, t.lab_color);
    db.build_similarities();
    return db;
}

// ----------------------------------------------------------- whole job
MaterialLibrary::MaterialLibrary(const std::vector<RgbImage>& materials, std::shared_ptr<Context> ctx)
    : ctx_(or_default(std::move(ctx))) {
    std::vector<const RgbImage*> images;
    for (const RgbImage& m : materials) images.push_back(&m);
    const PackedImages p = pack_images(images);
    check(mh_library_from_images(ctx_->get(), p.bytes.data(), p.offsets.data(), p.widths.data(),
                                 p.heights.data(), (int32_t)materials.size(), &lib_));
}

MaterialLibrary::~MaterialLibrary() {
    if (lib_) mh_library_destroy(lib_);
}

int32_t MaterialLibrary::size() const { return mh_library_size(lib_); }

MosaicOutput MaterialLibrary::generate_mosaic(const RgbImage& target,
                                              const MosaicSettings& s) const {
    mh_job_params p{};
    p.grid_w = s.grid_w;
    p.grid_h = s.grid_h;
    p.max_usage = s.max_usage_per_image;
    p.weight = s.adjacency_penalty_weight;
    p.color_strength = s.color_adjustment_strength;
    p.enable_optimization = s.enable_optimization ? 1 : 0;
    p.optimization_iterations = (int64_t)s.optimization_iterations;
    p.initial_temperature = s.optimization.initial_temperature;
    p.temperature_decay = s.optimization.temperature_decay;
    p.seed = s.optimization.seed;
    p.score_mode = (int32_t)s.score_mode;
    if (s.grid_w <= 0 || s.grid_h <= 0) throw Error(MH_ERR_INVALID_ARGUMENT, "grid must be positive");
    const uint32_t tw = target.width / (uint32_t)s.grid_w, th = target.height / (uint32_t)s.grid_h;
    MosaicOutput out;
    out.image = RgbImage(tw * (uint32_t)s.grid_w, th * (uint32_t)s.grid_h);
    out.grid.assign((size_t)s.grid_w * s.grid_h, MH_EMPTY);
    check(mh_generate(ctx_->get(), lib_, target.data.data(), (int32_t)target.width,
                      (int32_t)target.height, &p, out.grid.data(),
                      out.image.data.empty() ? nullptr : out.image.data.data(), &out.result));
    return out;
}

}  // namespace mosaic
