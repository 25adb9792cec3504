/*
 * mosaic_oracle.c -- CPU restatement of mosaic-art-rust's tile-matching path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Build with
 * -ffp-contract=off and without -ffast-math: the reference is Rust, whose f32
 * arithmetic is never contracted into FMAs.
 *
 * Third-party crates whose algorithms are restated here (none is vendored
 * under the reference; versions from Cargo.lock):
 *   palette 0.7.6      sRGB -> Lab (D65), sRGB <-> HSV
 *   kiddo 5.2.2        nearest_n / nearest_one (SquaredEuclidean)
 *   rand 0.8.5         thread_rng (unseeded) -- replaced by a counter RNG
 *   fast_image_resize 5.1.4  default Lanczos3 convolution
 */
#include "mosaic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Math building blocks                                                      */
/* ------------------------------------------------------------------------ */

/* sRGB transfer function on u8 inputs.  The reference feeds palette
 * `pixel as f32 / 255.0` (src/lib.rs:102-106); palette linearises with
 * c <= 0.04045 ? c/12.92 : ((c+0.055)/1.055)^2.4.  Only 256 inputs exist, so
 * the table is the definition: f32 c, linear branch in f32, power branch in
 * double rounded once to f32. */
void orc_srgb_lut(float out[256]) {
    for (int v = 0; v < 256; ++v) {
        float c = (float)v / 255.0f;
        if (c <= 0.04045f)
            out[v] = c / 12.92f;
        else
            out[v] = (float)pow(((double)c + 0.055) / 1.055, 2.4);
    }
}

/* Cube root: fdlibm-style exponent/3 seed (5 bits), then four Newton steps
 * t <- t - (t^3 - x)/(3 t^2) in IEEE double, rounded once to f32.  Only
 * +,-,*,/ are used so any IEEE machine reproduces the bits. */
float orc_cbrtf(float xf) {
    if (!(xf > 0.0f)) return xf; /* 0, negatives and NaN are never fed here */
    double x = (double)xf;
    uint64_t bits;
    memcpy(&bits, &x, 8);
    uint32_t hi = (uint32_t)(bits >> 32);
    hi = hi / 3u + 715094163u;
    uint64_t tb = (uint64_t)hi << 32;
    double t;
    memcpy(&t, &tb, 8);
    for (int i = 0; i < 4; ++i) {
        double t2 = t * t;
        t = t - (t2 * t - x) / (3.0 * t2);
    }
    return (float)t;
}

/* exp for f32 arguments: Cody-Waite reduction x = n ln2 + r (fdlibm split of
 * ln2), degree-13 Taylor polynomial for e^r in double (Horner), exact
 * scaling by 2^n, one rounding to f32. */
float orc_expf(float xf) {
    if (xf != xf) return xf;
    if (xf > 88.8f) return INFINITY;
    if (xf < -104.0f) return 0.0f;
    double x = (double)xf;
    double n = floor(x * 1.4426950408889634 + 0.5);
    double r = (x - n * 6.93147180369123816490e-01) - n * 1.90821492927058770002e-10;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int ni = (int)n;
    uint64_t sb = (uint64_t)(int64_t)(ni + 1023) << 52;
    double s;
    memcpy(&s, &sb, 8);
    return (float)(p * s);
}

/* ------------------------------------------------------------------------ */
/* a1: sRGB u8 -> Lab and the per-cell average                               */
/* ------------------------------------------------------------------------ */

static float g_lut[256];
static int g_lut_ready = 0;

static const float *lut(void) {
    if (!g_lut_ready) {
        orc_srgb_lut(g_lut);
        g_lut_ready = 1;
    }
    return g_lut;
}

static float clampf(float v, float lo, float hi) {
    /* Rust f32::clamp: NaN passes through */
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

static float lab_f(float t) {
    const float eps = (float)(216.0 / 24389.0);
    const float kappa = (float)(841.0 / 108.0);
    const float delta = (float)(4.0 / 29.0);
    if (t > eps) return orc_cbrtf(t);
    return kappa * t + delta;
}

/* palette Lab::from_color(Srgb<f32>) (called at src/lib.rs:107): linearise,
 * sRGB/D65 matrix (IEC 61966-2-1), divide by the D65 white point, CIE f(t),
 * L=116fy-16, a=500(fx-fy), b=200(fy-fz), then FromColor's clamp to
 * L in [0,100], a,b in [-128,127]. */
void orc_rgb8_to_lab(uint8_t r8, uint8_t g8, uint8_t b8, float out[3]) {
    const float *t = lut();
    float r = t[r8], g = t[g8], b = t[b8];
    float x = 0.4124564f * r + 0.3575761f * g + 0.1804375f * b;
    float y = 0.2126729f * r + 0.7151522f * g + 0.0721750f * b;
    float z = 0.0193339f * r + 0.1191920f * g + 0.9503041f * b;
    x = x / 0.95047f;
    y = y / 1.0f;
    z = z / 1.08883f;
    float fx = lab_f(x), fy = lab_f(y), fz = lab_f(z);
    float L = fy * 116.0f - 16.0f;
    float A = (fx - fy) * 500.0f;
    float B = (fy - fz) * 200.0f;
    out[0] = clampf(L, 0.0f, 100.0f);
    out[1] = clampf(A, -128.0f, 127.0f);
    out[2] = clampf(B, -128.0f, 127.0f);
}

/* MosaicGeneratorImpl::calculate_average_lab, src/lib.rs:94-119: serial f32
 * fold from (0,0,0) over pixels in row-major order, then each sum divided by
 * (w*h) as f32. */
void orc_average_lab(const uint8_t *rgb, int32_t w, int32_t h, int64_t row_stride, float out[3]) {
    float sl = 0.0f, sa = 0.0f, sb = 0.0f;
    for (int32_t y = 0; y < h; ++y) {
        const uint8_t *row = rgb + (int64_t)y * row_stride;
        for (int32_t x = 0; x < w; ++x) {
            float lab[3];
            orc_rgb8_to_lab(row[3 * x], row[3 * x + 1], row[3 * x + 2], lab);
            sl = sl + lab[0];
            sa = sa + lab[1];
            sb = sb + lab[2];
        }
    }
    float total = (float)((uint32_t)w * (uint32_t)h);
    out[0] = sl / total;
    out[1] = sa / total;
    out[2] = sb / total;
}

/* Per-cell average Lab of the target: src/main.rs:495-496 (tile size =
 * W/gw x H/gh), :529-541 (crop_imm at (x*tw, y*th) in raster order). */
void orc_cell_lab_means(const uint8_t *rgb, int32_t width, int32_t height, int32_t grid_w,
                        int32_t grid_h, float *out) {
    int32_t tw = width / grid_w, th = height / grid_h;
    int64_t stride = (int64_t)width * 3;
    for (int32_t y = 0; y < grid_h; ++y)
        for (int32_t x = 0; x < grid_w; ++x) {
            const uint8_t *base = rgb + (int64_t)(y * th) * stride + (int64_t)(x * tw) * 3;
            orc_average_lab(base, tw, th, stride, out + 3 * ((int64_t)y * grid_w + x));
        }
}

/* Per-material average Lab (material ingest, src/main.rs:263-301). */
void orc_stack_lab_means(const uint8_t *rgb, int32_t n, int32_t w, int32_t h, float *out) {
    for (int32_t i = 0; i < n; ++i)
        orc_average_lab(rgb + (int64_t)i * w * h * 3, w, h, (int64_t)w * 3, out + 3 * i);
}

/* ------------------------------------------------------------------------ */
/* a2: similarity database                                                   */
/* ------------------------------------------------------------------------ */

/* calculate_lab_distance, src/similarity.rs:141-146. */
float orc_lab_distance(const float *l1, const float *l2) {
    float dl = l1[0] - l2[0];
    float da = l1[1] - l2[1];
    float db = l1[2] - l2[2];
    return sqrtf(dl * dl + da * da + db * db);
}

/* get_similarity by index, src/similarity.rs:85-103: diagonal is 0.0, the
 * stored value is distance(lab[min], lab[max]) (built at :74-81). */
float orc_sim(const float *lab, int32_t i, int32_t j) {
    if (i == j) return 0.0f;
    int32_t a = i < j ? i : j, b = i < j ? j : i;
    return orc_lab_distance(lab + 3 * a, lab + 3 * b);
}

void orc_sim_full(const float *lab, int32_t n, float *out) {
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = 0; j < n; ++j) out[(int64_t)i * n + j] = orc_sim(lab, i, j);
}

/* Rows `rows[0..nrows)` of the full matrix (out = nrows x n), for sampled
 * checks of large libraries. */
void orc_sim_rows(const float *lab, int32_t n, const int32_t *rows, int32_t nrows, float *out) {
    for (int32_t r = 0; r < nrows; ++r)
        for (int32_t j = 0; j < n; ++j) out[(int64_t)r * n + j] = orc_sim(lab, rows[r], j);
}

/* build_similarities, src/similarity.rs:69-82: packed upper triangle in
 * (i, j>i) order; position i*n - i*(i+1)/2 + j - i - 1 (:100). */
void orc_sim_packed(const float *lab, int32_t n, float *out) {
    int64_t pos = 0;
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = i + 1; j < n; ++j) out[pos++] = orc_lab_distance(lab + 3 * i, lab + 3 * j);
}

/* ------------------------------------------------------------------------ */
/* a3: K nearest materials                                                   */
/* ------------------------------------------------------------------------ */

static float sqdist(const float *q, const float *m) {
    /* kiddo SquaredEuclidean: fold of (q_i - m_i)^2 from 0 (0 + x is exact). */
    float d0 = q[0] - m[0], d1 = q[1] - m[1], d2 = q[2] - m[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

static uint64_t fkey(float d, int32_t idx) {
    uint32_t bits;
    memcpy(&bits, &d, 4);
    return ((uint64_t)bits << 32) | (uint32_t)idx;
}

static void heap_sift_down(uint64_t *h, int32_t n, int32_t i) {
    for (;;) {
        int32_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < n && h[l] > h[m]) m = l;
        if (r < n && h[r] > h[m]) m = r;
        if (m == i) return;
        uint64_t t = h[i];
        h[i] = h[m];
        h[m] = t;
        i = m;
    }
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* kiddo nearest_n::<SquaredEuclidean>(q, k) as called at src/main.rs:348-352:
 * the k materials with the smallest squared distance, ascending.  Ties are
 * ordered by material index (the build's declared rule; kiddo's own tie
 * order is unpinned).  Squared distances are >= +0, so the f32 bit pattern
 * orders like the value and (bits << 32 | idx) is a total order. */
void orc_nearest_n(const float *lab, int32_t n, const float *query, int32_t k, int32_t *out_idx,
                   float *out_sqdist) {
    if (k > n) k = n;
    if (k <= 0) return;
    uint64_t *heap = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)k);
    int32_t hn = 0;
    for (int32_t i = 0; i < n; ++i) {
        uint64_t key = fkey(sqdist(query, lab + 3 * i), i);
        if (hn < k) {
            /* sift up */
            int32_t c = hn++;
            heap[c] = key;
            while (c > 0) {
                int32_t p = (c - 1) / 2;
                if (heap[p] >= heap[c]) break;
                uint64_t t = heap[p];
                heap[p] = heap[c];
                heap[c] = t;
                c = p;
            }
        } else if (key < heap[0]) {
            heap[0] = key;
            heap_sift_down(heap, hn, 0);
        }
    }
    qsort(heap, (size_t)hn, sizeof(uint64_t), cmp_u64);
    for (int32_t r = 0; r < hn; ++r) {
        uint32_t bits = (uint32_t)(heap[r] >> 32);
        out_idx[r] = (int32_t)(uint32_t)(heap[r] & 0xffffffffu);
        memcpy(&out_sqdist[r], &bits, 4);
    }
    free(heap);
}

/* The same K nearest from a k-d tree, as the reference finds them (kiddo's
 * KdTree<f32, u64, 3, 256, u32>, src/main.rs:145-151; buckets of 256).  The
 * answer equals orc_nearest_n's exactly, ties included:
 *   * every point's key is the same (sqdist, idx) from the same sqdist();
 *   * a far subtree is skipped only when fl((q_a - split)^2) is strictly
 *     greater than the current K-th smallest sqdist.  Rounding is monotone,
 *     so every point beyond the split plane has an f32 sqdist at least that
 *     bound, hence a key above the K-th smallest: skipping it loses nothing,
 *     and points at exactly the bound (possible ties) are still visited.
 * Nodes split at the median of their widest axis; points equal to the split
 * may sit on either side, which the bound above allows. */
typedef struct {
    int32_t axis; /* -1: leaf */
    float split;
    int32_t lo, hi;       /* children (inner) */
    int32_t start, count; /* points[start .. start+count) (leaf) */
} kd_node;

struct orc_kdtree {
    const float *lab;
    int32_t n;
    int32_t *idx;
    kd_node *nodes;
    int32_t n_nodes;
};

#define KD_BUCKET 256

static float kd_coord(const orc_kdtree *t, int32_t i, int32_t a) { return t->lab[3 * (int64_t)i + a]; }

/* Partial sort of idx[lo..hi) so position k holds its order statistic by
 * coordinate a (quickselect). */
static void kd_select(orc_kdtree *t, int32_t lo, int32_t hi, int32_t k, int32_t a) {
    int32_t *v = t->idx;
    while (hi - lo > 1) {
        float pivot = kd_coord(t, v[lo + (hi - lo) / 2], a);
        int32_t i = lo, j = hi - 1;
        while (i <= j) {
            while (kd_coord(t, v[i], a) < pivot) ++i;
            while (kd_coord(t, v[j], a) > pivot) --j;
            if (i <= j) {
                int32_t tmp = v[i];
                v[i] = v[j];
                v[j] = tmp;
                ++i;
                --j;
            }
        }
        if (k <= j)
            hi = j + 1;
        else if (k >= i)
            lo = i;
        else
            return;
    }
}

static int32_t kd_build_node(orc_kdtree *t, int32_t start, int32_t count) {
    int32_t id = t->n_nodes++;
    kd_node *nd = &t->nodes[id];
    nd->start = start;
    nd->count = count;
    nd->axis = -1;
    if (count <= KD_BUCKET) return id;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = start; i < start + count; ++i)
        for (int32_t a = 0; a < 3; ++a) {
            float c = kd_coord(t, t->idx[i], a);
            lo[a] = c < lo[a] ? c : lo[a];
            hi[a] = c > hi[a] ? c : hi[a];
        }
    int32_t axis = 0;
    for (int32_t a = 1; a < 3; ++a)
        if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
    if (!(hi[axis] > lo[axis])) return id; /* all points equal: one leaf */
    int32_t mid = start + count / 2;
    kd_select(t, start, start + count, mid, axis);
    float split = kd_coord(t, t->idx[mid], axis);
    int32_t l = kd_build_node(t, start, mid - start);
    int32_t r = kd_build_node(t, mid, start + count - mid);
    nd = &t->nodes[id]; /* (t->nodes does not move: sized up front) */
    nd->axis = axis;
    nd->split = split;
    nd->lo = l;
    nd->hi = r;
    return id;
}

orc_kdtree *orc_kdtree_build(const float *lab, int32_t n) {
    orc_kdtree *t = (orc_kdtree *)calloc(1, sizeof(orc_kdtree));
    t->lab = lab;
    t->n = n;
    t->idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; ++i) t->idx[i] = i;
    /* a binary tree over n points with leaves of at most KD_BUCKET points
     * has fewer than 2n/KD_BUCKET*2 + 1 nodes; 2n + 1 is ample */
    t->nodes = (kd_node *)malloc(sizeof(kd_node) * (size_t)(2 * (int64_t)n + 1));
    if (n > 0) kd_build_node(t, 0, n);
    return t;
}

void orc_kdtree_free(orc_kdtree *t) {
    if (!t) return;
    free(t->idx);
    free(t->nodes);
    free(t);
}

typedef struct {
    uint64_t *heap; /* max-heap of the best keys so far */
    int32_t hn, k;
} kd_query;

static void kd_offer(kd_query *q, uint64_t key) {
    if (q->hn < q->k) {
        int32_t c = q->hn++;
        q->heap[c] = key;
        while (c > 0) {
            int32_t p = (c - 1) / 2;
            if (q->heap[p] >= q->heap[c]) break;
            uint64_t t = q->heap[p];
            q->heap[p] = q->heap[c];
            q->heap[c] = t;
            c = p;
        }
    } else if (key < q->heap[0]) {
        q->heap[0] = key;
        heap_sift_down(q->heap, q->hn, 0);
    }
}

static void kd_search(const orc_kdtree *t, int32_t id, const float *query, kd_query *q) {
    const kd_node *nd = &t->nodes[id];
    if (nd->axis < 0) {
        for (int32_t i = nd->start; i < nd->start + nd->count; ++i) {
            int32_t m = t->idx[i];
            kd_offer(q, fkey(sqdist(query, t->lab + 3 * (int64_t)m), m));
        }
        return;
    }
    float d = query[nd->axis] - nd->split;
    int32_t near = d < 0.0f ? nd->lo : nd->hi, far = d < 0.0f ? nd->hi : nd->lo;
    kd_search(t, near, query, q);
    float bound = d * d;
    if (q->hn < q->k) {
        kd_search(t, far, query, q);
    } else {
        uint32_t wbits = (uint32_t)(q->heap[0] >> 32);
        float worst;
        memcpy(&worst, &wbits, 4);
        if (!(bound > worst)) kd_search(t, far, query, q);
    }
}

void orc_kdtree_nearest_n(const orc_kdtree *t, const float *query, int32_t k, int32_t *out_idx,
                          float *out_sqdist) {
    if (k > t->n) k = t->n;
    if (k <= 0) return;
    kd_query q;
    q.heap = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)k);
    q.hn = 0;
    q.k = k;
    kd_search(t, 0, query, &q);
    qsort(q.heap, (size_t)q.hn, sizeof(uint64_t), cmp_u64);
    for (int32_t r = 0; r < q.hn; ++r) {
        uint32_t bits = (uint32_t)(q.heap[r] >> 32);
        out_idx[r] = (int32_t)(uint32_t)(q.heap[r] & 0xffffffffu);
        memcpy(&out_sqdist[r], &bits, 4);
    }
    free(q.heap);
}

/* orc_kdtree_nearest_n for nq queries (out: nq x min(k, n)). */
void orc_nearest_n_kd(const float *lab, int32_t n, const float *queries, int32_t nq, int32_t k,
                      int32_t *out_idx, float *out_sqdist) {
    int32_t kk = k < n ? k : n;
    orc_kdtree *t = orc_kdtree_build(lab, n);
    for (int32_t i = 0; i < nq; ++i)
        orc_kdtree_nearest_n(t, queries + 3 * (int64_t)i, k, out_idx + (int64_t)i * kk,
                             out_sqdist + (int64_t)i * kk);
    orc_kdtree_free(t);
}

/* ------------------------------------------------------------------------ */
/* a4-a9: placement                                                          */
/* ------------------------------------------------------------------------ */

/* can_place_at_position, src/main.rs:309-333: no in-bounds 4-neighbour may
 * already hold the same material. */
static int can_place(const int32_t *grid, int32_t gw, int32_t gh, int32_t x, int32_t y, int32_t m) {
    if (x > 0 && grid[y * gw + x - 1] == m) return 0;
    if (x + 1 < gw && grid[y * gw + x + 1] == m) return 0;
    if (y > 0 && grid[(y - 1) * gw + x] == m) return 0;
    if (y + 1 < gh && grid[(y + 1) * gw + x] == m) return 0;
    return 1;
}

/* AdjacencyPenaltyCalculator::calculate_penalty, src/adjacency.rs:63-88 with
 * get_adjacent_positions order Up, Down, Left, Right (:17-45). */
static float calc_penalty(const float *lab, const int32_t *grid, int32_t gw, int32_t gh, int32_t x,
                          int32_t y, int32_t m, float weight) {
    float p = 0.0f;
    int32_t nb[4];
    int32_t cnt = 0;
    if (y > 0) nb[cnt++] = (y - 1) * gw + x;
    if (y < gh - 1) nb[cnt++] = (y + 1) * gw + x;
    if (x > 0) nb[cnt++] = y * gw + x - 1;
    if (x < gw - 1) nb[cnt++] = y * gw + x + 1;
    for (int32_t i = 0; i < cnt; ++i) {
        int32_t o = grid[nb[i]];
        if (o != ORC_EMPTY) p += 1.0f / (orc_sim(lab, m, o) + 1.0f);
    }
    return p * weight;
}

/* The placement loop of generate_mosaic (src/main.rs:529-545) driving
 * find_and_use_best_tile_with_position (:335-418) and
 * fallback_tile_selection (:420-475).  `order` lists cell indices y*gw+x in
 * processing order (NULL: raster order over the whole grid).  grid/usage
 * are in/out state (UsageTracker keyed by material index, src/lib.rs:57-89).
 * stage[cell]: 0 scored scan, 1 fallback first-placeable, 2 nearest_one,
 * 255 untouched.  dist/score: chosen candidate's colour distance and total
 * score (the fallbacks add no penalty).
 * score_mode 0 scores sqdist + penalty (src/main.rs:385-398); score_mode 1 is
 * the GUI generator's find_best_tile_primary (src/gui/app_full.rs:1396-1475):
 * sqrt(sqdist) + penalty, same candidates, bans, usage rule and fallbacks
 * (:1477-1563).  dist holds the distance the mode scores with. */
void orc_place(const float *lab, int32_t n, const float *cell_lab, int32_t gw, int32_t gh,
               int32_t max_usage, float weight, int32_t score_mode, const int32_t *order,
               int32_t n_order, int32_t *grid, uint32_t *usage, uint8_t *stage, float *dist,
               float *score, orc_place_stats *st) {
    orc_place_stats s;
    memset(&s, 0, sizeof(s));
    int32_t cells = order ? n_order : gw * gh;
    int32_t K = n < 100 ? n : 100;
    int32_t cidx[100];
    float cd[100];
    /* the K nearest from a k-d tree, as the reference's kiddo tree
     * (src/main.rs:145-151); exactly orc_nearest_n's answer */
    orc_kdtree *tree = orc_kdtree_build(lab, n);
    for (int32_t c = 0; c < cells; ++c) {
        int32_t cell = order ? order[c] : c;
        int32_t x = cell % gw, y = cell / gw;
        if (n == 0) { /* src/main.rs:342-345: no tiles -> None, cell stays empty */
            s.empty++;
            continue;
        }
        orc_kdtree_nearest_n(tree, cell_lab + 3 * (int64_t)cell, K, cidx, cd);
        if (score_mode == 1) /* app_full.rs:1436-1439: sqrt of the same f32 sum */
            for (int32_t r = 0; r < K; ++r) cd[r] = sqrtf(cd[r]);
        int32_t best = -1;
        float best_score = 0.0f;
        for (int32_t r = 0; r < K; ++r) {
            int32_t m = cidx[r];
            if (!(usage[m] < (uint32_t)max_usage)) continue; /* :375 */
            if (!can_place(grid, gw, gh, x, y, m)) continue;   /* :380 */
            float pen = weight > 0.0f ? calc_penalty(lab, grid, gw, gh, x, y, m, weight) : 0.0f;
            float total = cd[r] + pen;                           /* :398 */
            if (best < 0 || total < best_score) {                /* :401-407 */
                best = r;
                best_score = total;
            }
        }
        if (best >= 0) {
            int32_t m = cidx[best];
            usage[m] += 1;
            grid[cell] = m;
            if (stage) stage[cell] = 0;
            if (dist) dist[cell] = cd[best];
            if (score) score[cell] = best_score;
            s.main_path++;
            s.placed++;
            continue;
        }
        /* fallback_tile_selection: reset ALL usage counts (:433) */
        memset(usage, 0, sizeof(uint32_t) * (size_t)n);
        s.resets++;
        int32_t pick = -1;
        for (int32_t r = 0; r < K; ++r)
            if (can_place(grid, gw, gh, x, y, cidx[r])) { /* :441-453, no usage check */
                pick = r;
                break;
            }
        int32_t st_code = 1;
        if (pick < 0) { /* nearest_one (:456-474) = first key in (sqdist, idx) order */
            pick = 0;
            st_code = 2;
        }
        int32_t m = cidx[pick];
        usage[m] += 1;
        grid[cell] = m;
        if (stage) stage[cell] = (uint8_t)st_code;
        if (dist) dist[cell] = cd[pick];
        if (score) score[cell] = cd[pick];
        if (st_code == 1)
            s.fallback_ban++;
        else
            s.fallback_near++;
        s.placed++;
    }
    orc_kdtree_free(tree);
    if (st) *st = s;
}

/* ------------------------------------------------------------------------ */
/* a10-a12: cost, swap delta, annealing                                      */
/* ------------------------------------------------------------------------ */

static float term(const float *lab, int32_t a, int32_t b) {
    return 1.0f / (orc_sim(lab, a, b) + 1.0f);
}

/* calculate_total_cost, src/adjacency.rs:91-130: raster order; right then
 * down neighbour; unweighted. */
float orc_total_cost(const float *lab, int32_t n, const int32_t *grid, int32_t gw, int32_t gh) {
    (void)n;
    float total = 0.0f;
    if (gh == 0) return 0.0f;
    for (int32_t y = 0; y < gh; ++y)
        for (int32_t x = 0; x < gw; ++x) {
            int32_t c = grid[y * gw + x];
            if (c == ORC_EMPTY) continue;
            if (x < gw - 1) {
                int32_t r = grid[y * gw + x + 1];
                if (r != ORC_EMPTY) total += term(lab, c, r);
            }
            if (y < gh - 1) {
                int32_t d = grid[(y + 1) * gw + x];
                if (d != ORC_EMPTY) total += term(lab, c, d);
            }
        }
    return total;
}

static int32_t adjacent(int32_t gw, int32_t gh, int32_t x, int32_t y, int32_t *ax, int32_t *ay) {
    int32_t c = 0;
    if (y > 0) { ax[c] = x; ay[c] = y - 1; c++; }
    if (y < gh - 1) { ax[c] = x; ay[c] = y + 1; c++; }
    if (x > 0) { ax[c] = x - 1; ay[c] = y; c++; }
    if (x < gw - 1) { ax[c] = x + 1; ay[c] = y; c++; }
    return c;
}

/* calculate_swap_delta, src/adjacency.rs:133-214 (unweighted). */
float orc_swap_delta(const float *lab, int32_t n, const int32_t *grid, int32_t gw, int32_t gh,
                     int32_t x1, int32_t y1, int32_t x2, int32_t y2) {
    (void)n;
    if (gh == 0) return 0.0f;
    int32_t m1 = grid[y1 * gw + x1], m2 = grid[y2 * gw + x2];
    if (m1 == ORC_EMPTY || m2 == ORC_EMPTY) return 0.0f;
    if (m1 == m2) return 0.0f;
    float old_cost = 0.0f, new_cost = 0.0f;
    int32_t ax[4], ay[4];
    int32_t c1 = adjacent(gw, gh, x1, y1, ax, ay);
    int is_adj = 0;
    for (int32_t i = 0; i < c1; ++i) {
        if (ax[i] == x2 && ay[i] == y2) {
            is_adj = 1;
            continue;
        }
        int32_t a = grid[ay[i] * gw + ax[i]];
        if (a == ORC_EMPTY) continue;
        old_cost += term(lab, m1, a);
        new_cost += term(lab, m2, a);
    }
    int32_t c2 = adjacent(gw, gh, x2, y2, ax, ay);
    for (int32_t i = 0; i < c2; ++i) {
        if (ax[i] == x1 && ay[i] == y1) continue;
        int32_t a = grid[ay[i] * gw + ax[i]];
        if (a == ORC_EMPTY) continue;
        old_cost += term(lab, m2, a);
        new_cost += term(lab, m1, a);
    }
    if (is_adj) {
        float t = term(lab, m1, m2);
        old_cost += t;
        new_cost += t;
    }
    return new_cost - old_cost;
}

/* Counter-based draw k of proposal `it`: the SplitMix64 finaliser (Steele,
 * Lea, Flood 2014) of seed + golden*(8*it + k + 1), high 32 bits.  Replaces
 * rand::thread_rng (src/optimizer.rs:48), whose stream is unseeded. */
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint32_t orc_draw(uint64_t seed, uint64_t it, uint32_t k) {
    return (uint32_t)(mix64(seed + 0x9E3779B97F4A7C15ULL * (it * 8u + k + 1u)) >> 32);
}

static int32_t draw_range(uint64_t seed, uint64_t it, uint32_t k, int32_t bound) {
    return (int32_t)(((uint64_t)orc_draw(seed, it, k) * (uint64_t)(uint32_t)bound) >> 32);
}

static float draw_unit(uint64_t seed, uint64_t it) {
    return (float)(orc_draw(seed, it, 4) >> 8) * (1.0f / 16777216.0f);
}

/* MosaicOptimizer::optimize (src/optimizer.rs:41-129) and optimize_greedy
 * (:132-198).  Per iteration: x1,y1,x2,y2 uniform; same cell or an empty
 * cell -> skip WITHOUT cooling (:66-73); delta (:76); accept if delta < 0,
 * else (SA only) if U[0,1) < exp(-delta/T) (:79-84); on accept swap,
 * cost += delta, track best (:86-99); T *= decay (:102).  trace_delta[it]
 * gets the delta (NaN when skipped), trace_accept[it] 1/0. */
void orc_anneal(const float *lab, int32_t n, int32_t *grid, int32_t gw, int32_t gh,
                int64_t max_iterations, float t0, float decay, uint64_t seed, int32_t greedy,
                orc_anneal_result *res, float *trace_delta, uint8_t *trace_accept) {
    orc_anneal_result r;
    memset(&r, 0, sizeof(r));
    if (gh == 0 || gw == 0) { /* empty grid: OptimizationResult::default() */
        if (res) *res = r;
        return;
    }
    float cost = orc_total_cost(lab, n, grid, gw, gh);
    r.initial_cost = cost;
    r.best_cost = cost;
    float temp = t0;
    for (int64_t it = 0; it < max_iterations; ++it) {
        int32_t x1 = draw_range(seed, (uint64_t)it, 0, gw);
        int32_t y1 = draw_range(seed, (uint64_t)it, 1, gh);
        int32_t x2 = draw_range(seed, (uint64_t)it, 2, gw);
        int32_t y2 = draw_range(seed, (uint64_t)it, 3, gh);
        if (trace_delta) trace_delta[it] = NAN;
        if (trace_accept) trace_accept[it] = 0;
        if (x1 == x2 && y1 == y2) continue;
        if (grid[y1 * gw + x1] == ORC_EMPTY || grid[y2 * gw + x2] == ORC_EMPTY) continue;
        float delta = orc_swap_delta(lab, n, grid, gw, gh, x1, y1, x2, y2);
        r.evaluated++;
        int accept;
        if (greedy) {
            accept = delta < 0.0f;
        } else if (delta < 0.0f) {
            accept = 1;
        } else {
            float prob = orc_expf(-delta / temp);
            accept = draw_unit(seed, (uint64_t)it) < prob;
        }
        if (trace_delta) trace_delta[it] = delta;
        if (accept) {
            int32_t t = grid[y1 * gw + x1];
            grid[y1 * gw + x1] = grid[y2 * gw + x2];
            grid[y2 * gw + x2] = t;
            cost += delta;
            r.accepted_count++;
            if (greedy) {
                r.improved_count++;
            } else if (cost < r.best_cost) {
                r.best_cost = cost;
                r.improved_count++;
            }
            if (trace_accept) trace_accept[it] = 1;
        }
        if (!greedy) temp *= decay;
    }
    r.final_cost = cost;
    if (greedy) r.best_cost = cost; /* :193 */
    r.iterations = max_iterations;
    if (res) *res = r;
}

/* ------------------------------------------------------------------------ */
/* a13-a16: average RGB, colour adjustment, resize, composite                */
/* ------------------------------------------------------------------------ */

/* calculate_average_rgb, src/main.rs:687-708: u64 sums, truncating divide. */
void orc_average_rgb(const uint8_t *rgb, int32_t w, int32_t h, int64_t row_stride, uint8_t out[3]) {
    uint64_t s[3] = {0, 0, 0};
    for (int32_t y = 0; y < h; ++y)
        for (int32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) s[c] += rgb[(int64_t)y * row_stride + 3 * x + c];
    uint64_t cnt = (uint64_t)w * (uint64_t)h;
    for (int c = 0; c < 3; ++c) out[c] = (uint8_t)(s[c] / cnt);
}

/* palette Hsv::from_color(Srgb<f32>) (used at src/color_adjustment.rs:70,
 * :122-123): max/min/sep/coeff form, hue = (sep/d + coeff)*60 (may be
 * negative; RgbHue keeps the raw value), s = d/max, v = max. */
void orc_rgb_to_hsv(float r, float g, float b, float out[3]) {
    r = r > 0.0f ? r : 0.0f;
    g = g > 0.0f ? g : 0.0f;
    b = b > 0.0f ? b : 0.0f;
    float mx, mn, sep, coeff;
    if (r > g) {
        mx = r; mn = g; sep = g - b; coeff = 0.0f;
    } else {
        mx = g; mn = r; sep = b - r; coeff = 2.0f;
    }
    if (b > mx) {
        mx = b; sep = r - g; coeff = 4.0f;
    } else if (b < mn) {
        mn = b;
    }
    float h = 0.0f, s = 0.0f;
    if (mx != mn) {
        float d = mx - mn;
        h = (sep / d + coeff) * 60.0f;
        s = d / mx;
    }
    out[0] = h;
    out[1] = clampf(s, 0.0f, 1.0f);
    out[2] = clampf(mx, 0.0f, 1.0f);
}

/* palette Srgb::from_color(Hsv) (src/color_adjustment.rs:76): hue
 * normalised to [0,360) by deg - floor(deg/360)*360, sector form, then
 * FromColor's clamp to [0,1]. */
void orc_hsv_to_rgb(float hdeg, float s, float v, float out[3]) {
    float c = v * s;
    float hp = hdeg - floorf(hdeg / 360.0f) * 360.0f;
    float h = hp / 60.0f;
    float hm2 = h - 2.0f * floorf(h * 0.5f); /* == fmodf(h, 2) for h >= 0 */
    float x = c * (1.0f - fabsf(hm2 - 1.0f));
    float m = v - c;
    float r, g, b;
    if (h >= 0.0f && h < 1.0f) { r = c; g = x; b = 0.0f; }
    else if (h >= 1.0f && h < 2.0f) { r = x; g = c; b = 0.0f; }
    else if (h >= 2.0f && h < 3.0f) { r = 0.0f; g = c; b = x; }
    else if (h >= 3.0f && h < 4.0f) { r = 0.0f; g = x; b = c; }
    else if (h >= 4.0f && h < 5.0f) { r = x; g = 0.0f; b = c; }
    else { r = c; g = 0.0f; b = x; }
    out[0] = clampf(r + m, 0.0f, 1.0f);
    out[1] = clampf(g + m, 0.0f, 1.0f);
    out[2] = clampf(b + m, 0.0f, 1.0f);
}

/* ColorAdjustment::new, src/color_adjustment.rs:30-37 (clamps). adj =
 * {brightness, contrast, hue_shift, saturation}. */
void orc_adjustment_new(float brightness, float contrast, float hue_shift, float saturation,
                        float out[4]) {
    out[0] = clampf(brightness, -1.0f, 1.0f);
    out[1] = clampf(contrast, 0.0f, 2.0f);
    out[2] = clampf(hue_shift, -180.0f, 180.0f);
    out[3] = clampf(saturation, 0.0f, 2.0f);
}

/* calculate_optimal_adjustment, src/color_adjustment.rs:99-154. */
void orc_optimal_adjustment(const uint8_t tile[3], const uint8_t target[3], float k, float out[4]) {
    float tr = tile[0] / 255.0f, tg = tile[1] / 255.0f, tb = tile[2] / 255.0f;
    float gr = target[0] / 255.0f, gg = target[1] / 255.0f, gb = target[2] / 255.0f;
    float tile_luma = 0.299f * tr + 0.587f * tg + 0.114f * tb;
    float target_luma = 0.299f * gr + 0.587f * gg + 0.114f * gb;
    float brightness = (target_luma - tile_luma) * k;
    float th[3], gh[3];
    orc_rgb_to_hsv(tr, tg, tb, th);
    orc_rgb_to_hsv(gr, gg, gb, gh);
    float hue = 0.0f;
    if (th[1] > 0.1f && gh[1] > 0.1f) {
        float diff = gh[0] - th[0];
        float wrapped = diff > 180.0f ? diff - 360.0f : (diff < -180.0f ? diff + 360.0f : diff);
        hue = wrapped * k * 0.5f;
    }
    float sat = 1.0f;
    if (th[1] > 0.01f) {
        float ratio = gh[1] / th[1];
        sat = 1.0f + (ratio - 1.0f) * k * 0.7f;
    }
    orc_adjustment_new(brightness, 1.0f, hue, sat, out);
}

static float bc(float v, float brightness, float contrast) {
    /* apply_brightness_contrast, src/color_adjustment.rs:91-96 */
    float c = clampf((v - 0.5f) * contrast + 0.5f, 0.0f, 1.0f);
    return clampf(c + brightness, 0.0f, 1.0f);
}

static uint8_t to_u8(float v) {
    /* (v * 255.0).clamp(0.0, 255.0) as u8: saturating truncation, NaN -> 0 */
    float t = clampf(v * 255.0f, 0.0f, 255.0f);
    if (t != t) return 0;
    return (uint8_t)t;
}

/* ColorAdjustment::adjust_pixel, src/color_adjustment.rs:53-87. */
void orc_adjust_pixel(const float adj[4], const uint8_t in[3], uint8_t out[3]) {
    float r = in[0] / 255.0f, g = in[1] / 255.0f, b = in[2] / 255.0f;
    float br = bc(r, adj[0], adj[1]), bg = bc(g, adj[0], adj[1]), bb = bc(b, adj[0], adj[1]);
    float fr = br, fg = bg, fb = bb;
    if (adj[2] != 0.0f || adj[3] != 1.0f) {
        float hsv[3], rgb[3];
        orc_rgb_to_hsv(br, bg, bb, hsv);
        float h = hsv[0] + adj[2];
        float hr = fmodf(h, 360.0f); /* rem_euclid(360.0) */
        if (hr < 0.0f) hr = hr + 360.0f;
        float s = clampf(hsv[1] * adj[3], 0.0f, 1.0f);
        orc_hsv_to_rgb(hr, s, hsv[2], rgb);
        fr = rgb[0];
        fg = rgb[1];
        fb = rgb[2];
    }
    out[0] = to_u8(fr);
    out[1] = to_u8(fg);
    out[2] = to_u8(fb);
}

/* ColorAdjustment::apply_to_image, src/color_adjustment.rs:40-50. */
void orc_apply_adjustment(const float adj[4], const uint8_t *in, int64_t npx, uint8_t *out) {
    for (int64_t i = 0; i < npx; ++i) orc_adjust_pixel(adj, in + 3 * i, out + 3 * i);
}

/* Lanczos3 resampling, fast_image_resize 5.1.4's default Convolution(Lanczos3)
 * for u8 pixels (src/main.rs:710-738; the crate follows Pillow-SIMD):
 *  - scale = in/out, fs = max(scale, 1), radius r = 3 fs;
 *  - output o centres on c = (o + 0.5) scale and reads source pixels
 *    [max(floor(c - r), 0), min(ceil(c + r), in)); pixel x has the weight
 *    lanczos3((x - (c - 0.5)) * (1 / fs)), and the weights are divided by
 *    their sum;
 *  - one precision p per axis: stepping p = 0, 1, ..., 21, stop at the first p
 *    for which round(max weight * 2^(p + 1)) >= 2^15; weights are rounded
 *    half away from zero to p fraction bits (16-bit integers);
 *  - a pass sums (1 << (p - 1)) + pixel * weight in 32 bits and keeps
 *    (sum >> p) clamped to 0..255; horizontal pass to u8, then vertical. */
static double sinc(double x) {
    if (x == 0.0) return 1.0;
    x = x * M_PI;
    return sin(x) / x;
}

static double lanczos3(double x) {
    if (x > -3.0 && x < 3.0) return sinc(x) * sinc(x / 3.0);
    return 0.0;
}

static int32_t lz_taps(int32_t in_size, int32_t out_size, int32_t *lo, int32_t *cnt, int32_t **kk,
                       int32_t *prec_out) {
    double scale = (double)in_size / (double)out_size;
    double fs = scale < 1.0 ? 1.0 : scale;
    double radius = 3.0 * fs;
    double recip = 1.0 / fs;
    int32_t ksize = (int32_t)ceil(radius) * 2 + 1;
    int32_t *k = (int32_t *)calloc((size_t)out_size * (size_t)ksize, sizeof(int32_t));
    double *w = (double *)calloc((size_t)out_size * (size_t)ksize, sizeof(double));
    double max_w = 0.0;
    for (int32_t o = 0; o < out_size; ++o) {
        double in_center = ((double)o + 0.5) * scale;
        double fmin = floor(in_center - radius), fmax = ceil(in_center + radius);
        int32_t xmin = fmin > 0.0 ? (int32_t)fmin : 0;
        int32_t xmax = fmax < (double)in_size ? (int32_t)fmax : in_size;
        double center = in_center - 0.5;
        double *wo = w + (int64_t)o * ksize;
        double ww = 0.0;
        for (int32_t x = xmin; x < xmax; ++x) {
            wo[x - xmin] = lanczos3(((double)x - center) * recip);
            ww += wo[x - xmin];
        }
        if (ww != 0.0)
            for (int32_t t = 0; t < xmax - xmin; ++t) wo[t] /= ww;
        for (int32_t t = 0; t < xmax - xmin; ++t)
            if (wo[t] > max_w) max_w = wo[t];
        lo[o] = xmin;
        cnt[o] = xmax - xmin;
    }
    int32_t prec = 0;
    for (int32_t p = 0; p < 22; ++p) {
        prec = p;
        if (round(max_w * (double)(1 << (p + 1))) >= 32768.0) break;
    }
    for (int64_t i = 0; i < (int64_t)out_size * ksize; ++i)
        k[i] = (int32_t)round(w[i] * (double)(1 << prec)); /* half away from zero */
    free(w);
    *kk = k;
    *prec_out = prec;
    return ksize;
}

static uint8_t clip8(int32_t acc, int32_t prec) {
    int32_t v = acc >> prec;
    return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

void orc_resize_lanczos3(const uint8_t *src, int32_t sw, int32_t sh, uint8_t *dst, int32_t dw,
                         int32_t dh) {
    if (dw <= 0 || dh <= 0) return;
    int32_t *xlo = (int32_t *)malloc(sizeof(int32_t) * (size_t)dw);
    int32_t *xc = (int32_t *)malloc(sizeof(int32_t) * (size_t)dw);
    int32_t *ylo = (int32_t *)malloc(sizeof(int32_t) * (size_t)dh);
    int32_t *yc = (int32_t *)malloc(sizeof(int32_t) * (size_t)dh);
    int32_t *kx, *ky, px, py;
    int32_t ksx = lz_taps(sw, dw, xlo, xc, &kx, &px);
    int32_t ksy = lz_taps(sh, dh, ylo, yc, &ky, &py);
    uint8_t *tmp = (uint8_t *)malloc((size_t)sh * (size_t)dw * 3);
    for (int32_t y = 0; y < sh; ++y)
        for (int32_t o = 0; o < dw; ++o)
            for (int c = 0; c < 3; ++c) {
                int32_t acc = 1 << (px - 1);
                for (int32_t t = 0; t < xc[o]; ++t)
                    acc += (int32_t)src[((int64_t)y * sw + xlo[o] + t) * 3 + c] * kx[(int64_t)o * ksx + t];
                tmp[((int64_t)y * dw + o) * 3 + c] = clip8(acc, px);
            }
    for (int32_t o = 0; o < dh; ++o)
        for (int32_t x = 0; x < dw; ++x)
            for (int c = 0; c < 3; ++c) {
                int32_t acc = 1 << (py - 1);
                for (int32_t t = 0; t < yc[o]; ++t)
                    acc += (int32_t)tmp[((int64_t)(ylo[o] + t) * dw + x) * 3 + c] * ky[(int64_t)o * ksy + t];
                dst[((int64_t)o * dw + x) * 3 + c] = clip8(acc, py);
            }
    free(tmp);
    free(kx);
    free(ky);
    free(xlo);
    free(xc);
    free(ylo);
    free(yc);
}

/* Compositing, src/main.rs:548-565 + :583-590 (and the post-SA rebuild
 * :629-661): per filled cell, resize the material to tw x th, and when
 * strength > 0 adjust it towards the target cell's average RGB; write at
 * (x*tw, y*th) into a gw*tw x gh*th image (zero elsewhere, :515-517). */
/* One cell of the composite (src/main.rs:548-590, :625-664): the resized
 * tile, colour-adjusted toward the target cell's average RGB when strength
 * > 0, copied to (x*tw, y*th).  adj_buf holds tw*th*3 bytes. */
static void composite_cell(const uint8_t *target, int64_t tstride, const uint8_t *tile, int32_t x,
                           int32_t y, int32_t tw, int32_t th, int32_t ow, float strength,
                           uint8_t *adj_buf, uint8_t *out) {
    const uint8_t *src = tile;
    if (strength > 0.0f) {
        uint8_t tavg[3], savg[3];
        float a[4];
        orc_average_rgb(target + (int64_t)(y * th) * tstride + (int64_t)(x * tw) * 3, tw, th,
                        tstride, tavg);
        orc_average_rgb(tile, tw, th, (int64_t)tw * 3, savg);
        orc_optimal_adjustment(savg, tavg, strength, a);
        orc_apply_adjustment(a, tile, (int64_t)tw * th, adj_buf);
        src = adj_buf;
    }
    for (int32_t dy = 0; dy < th; ++dy)
        memcpy(out + ((int64_t)(y * th + dy) * ow + (int64_t)x * tw) * 3,
               src + (int64_t)dy * tw * 3, (size_t)tw * 3);
}

void orc_composite(const uint8_t *target, int32_t width, int32_t height, const uint8_t *materials,
                   int32_t n, int32_t mw, int32_t mh, const int32_t *grid, int32_t gw, int32_t gh,
                   float strength, uint8_t *out) {
    int32_t tw = width / gw, th = height / gh;
    int32_t ow = gw * tw, oh = gh * th;
    memset(out, 0, (size_t)ow * (size_t)oh * 3);
    if (tw <= 0 || th <= 0) return;
    size_t tile_bytes = (size_t)tw * (size_t)th * 3;
    uint8_t *cache = (uint8_t *)malloc(tile_bytes * (size_t)(n > 0 ? n : 1));
    uint8_t *have = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    uint8_t *adj = (uint8_t *)malloc(tile_bytes);
    int64_t tstride = (int64_t)width * 3;
    for (int32_t y = 0; y < gh; ++y)
        for (int32_t x = 0; x < gw; ++x) {
            int32_t m = grid[y * gw + x];
            if (m == ORC_EMPTY) continue;
            uint8_t *tile = cache + tile_bytes * (size_t)m;
            if (!have[m]) {
                orc_resize_lanczos3(materials + (int64_t)m * mw * mh * 3, mw, mh, tile, tw, th);
                have[m] = 1;
            }
            composite_cell(target, tstride, tile, x, y, tw, th, ow, strength, adj, out);
        }
    free(cache);
    free(have);
    free(adj);
}

/* The composite from tiles already resized to the cell size (n x th x tw x 3,
 * tw = width/gw, th = height/gh): a library prepared once, as a resident
 * library's cached tiles are. */
void orc_composite_tiles(const uint8_t *target, int32_t width, int32_t height,
                         const uint8_t *tiles, const int32_t *grid, int32_t gw, int32_t gh,
                         float strength, uint8_t *out) {
    int32_t tw = width / gw, th = height / gh;
    int32_t ow = gw * tw, oh = gh * th;
    memset(out, 0, (size_t)ow * (size_t)oh * 3);
    if (tw <= 0 || th <= 0) return;
    size_t tile_bytes = (size_t)tw * (size_t)th * 3;
    uint8_t *adj = (uint8_t *)malloc(tile_bytes);
    int64_t tstride = (int64_t)width * 3;
    for (int32_t y = 0; y < gh; ++y)
        for (int32_t x = 0; x < gw; ++x) {
            int32_t m = grid[y * gw + x];
            if (m == ORC_EMPTY) continue;
            composite_cell(target, tstride, tiles + tile_bytes * (size_t)m, x, y, tw, th, ow,
                           strength, adj, out);
        }
    free(adj);
}
