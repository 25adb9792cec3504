"""Host checks of the exact shortcuts the device colour code takes
(mosaic-art-rust_amd/csrc/mh_device.hpp: u8_unit, fmod_360, hsv_to_rgb's
hue normalisation and its division by 60, rgb8_to_lab's divisions by the
white point).  Each shortcut is compared bit for bit with the IEEE
operation it replaces, over every input that matters (all 256 u8 values) or
over dense sweeps around the branch boundaries.  The C program below uses the
same f32 operations as the device helpers (fused multiply-add included); the
GPU parity tests then compare the device results with the oracle.  CPU only.
"""
import shutil
import subprocess

import pytest

PROG = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float u8_unit(uint32_t x) {
    const float a = (float)x;
    const float c = 1.0f / 255.0f;
    const float q = a * c;
    const float r = fmaf(-q, 255.0f, a);
    return fmaf(r, c, q);
}

static float fmod_360(float a) {
    const float m = fabsf(a);
    if (!(m < 720.0f)) return fmodf(a, 360.0f);
    const float r = m >= 360.0f ? m - 360.0f : m;
    return copysignf(r, a);
}

static float hue_norm_fast(float h) {
    return h >= 0.0f && h < 360.0f ? h + 0.0f : h - floorf(h / 360.0f) * 360.0f;
}
static float hue_norm(float h) { return h - floorf(h / 360.0f) * 360.0f; }

static float div60_fast(float x) {
    const float c = 1.0f / 60.0f;
    const float q = x * c;
    const float r = fmaf(-q, 60.0f, x);
    return fmaf(r, c, q);
}

int main(void) {
    long bad = 0, n = 0;
    for (uint32_t x = 0; x < 256; ++x, ++n)
        if (bits(u8_unit(x)) != bits((float)x / 255.0f)) { ++bad; printf("u8 %u\n", x); }
    /* every f32 in [-1000, 1000] with a stride, plus the neighbours of each
       boundary, signed zeros and non-finite values */
    float special[] = {0.0f, -0.0f, 360.0f, -360.0f, 720.0f, -720.0f, 1080.0f, -1080.0f,
                       INFINITY, -INFINITY, NAN, 1e30f, -1e30f, 1e-30f, -1e-30f};
    for (unsigned i = 0; i < sizeof special / sizeof *special; ++i) {
        float s = special[i];
        float c[3] = {s, nextafterf(s, INFINITY), nextafterf(s, -INFINITY)};
        for (int k = 0; k < 3; ++k, ++n) {
            float a = c[k];
            const float f = fmod_360(a), g = fmodf(a, 360.0f);
            if (isnan(g) ? !isnan(f) : bits(f) != bits(g)) { ++bad; printf("fmod %a\n", a); }
            const float u = hue_norm_fast(a), v = hue_norm(a);
            if (isnan(v) ? !isnan(u) : bits(u) != bits(v)) { ++bad; printf("hue %a\n", a); }
        }
    }
    /* [-1000, 1000] in steps of 2^-10, then every f32 within 2^-6 of each
       branch boundary +-360, +-720, and a grid around 0 */
    for (long i = -1024000; i <= 1024000; ++i, ++n) {
        const float a = (float)i * 0.0009765625f;
        if (bits(fmod_360(a)) != bits(fmodf(a, 360.0f))) { ++bad; if (bad < 10) printf("fmod %a\n", a); }
        if (bits(hue_norm_fast(a)) != bits(hue_norm(a))) { ++bad; if (bad < 10) printf("hue %a\n", a); }
    }
    const float edges[] = {360.0f, -360.0f, 720.0f, -720.0f};
    for (int e = 0; e < 4; ++e)
        for (float a = edges[e] - 0.015625f; a <= edges[e] + 0.015625f; a = nextafterf(a, INFINITY), ++n) {
            if (bits(fmod_360(a)) != bits(fmodf(a, 360.0f))) { ++bad; if (bad < 10) printf("fmod %a\n", a); }
            if (bits(hue_norm_fast(a)) != bits(hue_norm(a))) { ++bad; if (bad < 10) printf("hue %a\n", a); }
        }
    for (long i = -65536; i <= 65536; ++i, ++n) {  /* within 2^-6 of zero, steps of 2^-22 */
        const float a = (float)i * 2.384185791015625e-07f;
        if (bits(fmod_360(a)) != bits(fmodf(a, 360.0f))) { ++bad; if (bad < 10) printf("fmod %a\n", a); }
        if (bits(hue_norm_fast(a)) != bits(hue_norm(a))) { ++bad; if (bad < 10) printf("hue %a\n", a); }
    }
    /* the top of the fast hue range: every f32 in [359, 360) */
    for (float a = 359.0f; a < 360.0f; a = nextafterf(a, INFINITY), ++n)
        if (bits(hue_norm_fast(a)) != bits(hue_norm(a))) { ++bad; if (bad < 10) printf("hue %a\n", a); }
    /* div60_fast: every f32 in [2^-124, 360) and 0 */
    long bad60 = bits(div60_fast(0.0f)) != bits(0.0f / 60.0f), n60 = 1;
    const uint32_t lo = bits(0x1p-124f), hi = bits(360.0f);
    #pragma omp parallel for reduction(+ : bad60, n60) schedule(static)
    for (long u = lo; u < (long)hi; ++u) {
        float x;
        const uint32_t w = (uint32_t)u;
        memcpy(&x, &w, 4);
        if (bits(div60_fast(x)) != bits(x / 60.0f)) ++bad60;
        ++n60;
    }
    if (bad60) printf("div60 mismatches %ld\n", bad60);
    n += n60;
    bad += bad60;
    /* rgb8_to_lab: x / 0.95047f and z / 1.08883f through the rounded
       reciprocal and one fused correction, for the x and z of every u8
       colour (the linearised ramp as the context builds it) */
    float lut[256];
    for (int v = 0; v < 256; ++v) {
        const float c = (float)v / 255.0f;
        lut[v] = c <= 0.04045f ? c / 12.92f : (float)pow(((double)c + 0.055) / 1.055, 2.4);
    }
    const float rx = 1.0f / 0.95047f, rz = 1.0f / 1.08883f;
    long badw = 0;
    for (int r8 = 0; r8 < 256; ++r8)
        for (int g8 = 0; g8 < 256; ++g8)
            for (int b8 = 0; b8 < 256; ++b8, n += 2) {
                const float r = lut[r8], g = lut[g8], b = lut[b8];
                const float x = 0.4124564f * r + 0.3575761f * g + 0.1804375f * b;
                const float z = 0.0193339f * r + 0.1191920f * g + 0.9503041f * b;
                const float qx = x * rx, qz = z * rz;
                if (bits(fmaf(fmaf(-qx, 0.95047f, x), rx, qx)) != bits(x / 0.95047f)) ++badw;
                if (bits(fmaf(fmaf(-qz, 1.08883f, z), rz, qz)) != bits(z / 1.08883f)) ++badw;
            }
    if (badw) printf("white point mismatches %ld\n", badw);
    bad += badw;
    printf("checked %ld bad %ld\n", n, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_colour_adjustment_shortcuts_are_exact(tmp_path):
    src = tmp_path / "ident.c"
    src.write_text(PROG)
    exe = tmp_path / "ident"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", str(src), "-o",
                    str(exe), "-lm"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "bad 0" in r.stdout
