// Host-side check of the FP32 storage's colour word (lbm_frac.h): built with
// g++ by tests/test_frac_word.py, with the HIP qualifiers defined away.
#define __host__
#define __device__
#include "lbm_frac.h"

#include <cstdio>
#include <random>

using namespace psx_lbm_dev;

int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    // 1. fractions in [2^-126, 1/2]: rounding to nearest at 25 bits
    double max_rel = 0.0;
    long not_nearest = 0;
    for (long i = 0; i < 4000000; i++) {
        const double r = std::ldexp(0.5 + 0.5 * u(g), -1 - (int)(125.0 * u(g)));  // [2^-126, 1/2)
        const uint32_t w = frac_word(r);
        const double a = frac_value(w);
        max_rel = std::fmax(max_rel, std::fabs(a - r) / r);
        // the neighbours one unit of the last place away are no nearer
        const double up = frac_value(w + 2u), dn = w >= 2u ? frac_value(w - 2u) : 0.0;
        if (std::fabs(up - r) < std::fabs(a - r) || std::fabs(dn - r) < std::fabs(a - r)) not_nearest++;
    }
    // 2. tiny fractions below 2^-126: absolute error
    double max_abs_tiny = 0.0;
    for (long i = 0; i < 1000000; i++) {
        const double r = std::pow(2.0, -126.0 - 40.0 * u(g));
        max_abs_tiny = std::fmax(max_abs_tiny, std::fabs(frac_value(frac_word(r)) - r));
    }
    // 3. every word of the first 2^31 patterns (bit 0 clear) decodes and re-encodes to itself
    long roundtrip_bad = 0, words = 0;
    for (uint64_t w = 0; w < (126ull << 25) + (1ull << 25); w += 2u * 977u) {
        words++;
        if (frac_word(frac_value((uint32_t)w)) != (uint32_t)w) roundtrip_bad++;
    }
    // 4. populations: minority part's relative error, majority's error relative to f
    double max_min_rel = 0.0, max_maj = 0.0;
    for (long i = 0; i < 2000000; i++) {
        const double f = 0.01 + u(g);
        const double e = std::pow(10.0, -30.0 * u(g));
        double fr, fb;
        if (i % 2) { fr = f * e * 0.5; fb = f - fr; } else { fb = f * e * 0.5; fr = f - fb; }
        double a, b;
        frac_split(fr + fb, frac_encode(fr, fb), a, b);
        const double mn = std::fmin(fr, fb);
        const double emin = fr <= fb ? std::fabs(a - fr) : std::fabs(b - fb);
        max_min_rel = std::fmax(max_min_rel, emin / mn);
        max_maj = std::fmax(max_maj, std::fmax(std::fabs(a - fr), std::fabs(b - fb)) / f);
    }
    // 5. parts of opposite sign and fractions rounding to 1 (the escape)
    double max_esc = 0.0;
    long esc_words = 0;
    const double cases[][2] = {{-0.1, 1.0}, {1.2, -0.1}, {0.3, -0.29}, {-1e-3, 0.5}, {2.0, -1.9999}, {1e-30, -5e-31}};
    for (const auto &c : cases) {
        double a, b;
        float s = frac_encode(c[0], c[1]);
        uint32_t w;
        std::memcpy(&w, &s, 4);
        if ((w >> 25) == 127u) esc_words++;
        frac_split(c[0] + c[1], s, a, b);
        const double mn = std::fabs(c[0]) < std::fabs(c[1]) ? c[0] : c[1];
        const double got = std::fabs(c[0]) < std::fabs(c[1]) ? a : b;
        max_esc = std::fmax(max_esc, std::fabs(got - mn) / std::fabs(mn));
    }
    // 6. the extended form: 33-bit significand, half a unit of the last place
    double max_rel_x = 0.0, max_half_ulp_x = 0.0, max_abs_tiny_x = 0.0, max_min_rel_x = 0.0;
    long roundtrip_bad_x = 0, words_x = 0, red_x_nonzero = 0;
    for (long i = 0; i < 4000000; i++) {
        const int k = 1 + (int)(125.0 * u(g));
        const double r = std::ldexp(0.5 + 0.5 * u(g), -k);  // [2^-126, 1/2)
        uint32_t x;
        const uint32_t w = frac_word_t<8>(r, x);
        const double a = frac_value(w, x);
        max_rel_x = std::fmax(max_rel_x, std::fabs(a - r) / r);
        max_half_ulp_x = std::fmax(max_half_ulp_x, std::fabs(a - r) / std::ldexp(1.0, -k - 1 - 32));
    }
    for (long i = 0; i < 1000000; i++) {
        const double r = std::pow(2.0, -126.0 - 40.0 * u(g));
        uint32_t x;
        const uint32_t w = frac_word_t<8>(r, x);
        max_abs_tiny_x = std::fmax(max_abs_tiny_x, std::fabs(frac_value(w, x) - r));
    }
    for (uint64_t w = 0; w < (126ull << 25) + (1ull << 25); w += 2u * 7919u) {
        for (uint32_t x = 0; x < 256; x += 51) {
            words_x++;
            uint32_t x2;
            const uint32_t w2 = frac_word_t<8>(frac_value((uint32_t)w, x), x2);
            if (w2 != (uint32_t)w || x2 != x) roundtrip_bad_x++;
        }
    }
    for (long i = 0; i < 2000000; i++) {
        const double f = 0.01 + u(g);
        const double e = std::pow(10.0, -30.0 * u(g));
        double fr, fb;
        if (i % 2) { fr = f * e * 0.5; fb = f - fr; } else { fb = f * e * 0.5; fr = f - fb; }
        uint32_t x;
        const float s = frac_encode_t<8>(fr, fb, x);
        double a, b;
        frac_split(fr + fb, s, a, b, x);
        const double mn = std::fmin(fr, fb);
        const double emin = fr <= fb ? std::fabs(a - fr) : std::fabs(b - fb);
        max_min_rel_x = std::fmax(max_min_rel_x, emin / mn);
    }
    {
        uint32_t x;
        frac_encode_t<8>(1.0, 0.0, x);
        red_x_nonzero += x != 0;
        frac_encode_t<8>(0.0, 1.0, x);
        red_x_nonzero += x != 0;
    }
    printf("{\"max_rel_x\": %.6e, \"max_half_ulp_x\": %.6f, \"max_abs_tiny_x\": %.6e, \"roundtrip_words_x\": %ld, "
           "\"roundtrip_bad_x\": %ld, \"max_min_rel_x\": %.6e, \"pure_x_nonzero\": %ld}\n",
           max_rel_x, max_half_ulp_x, max_abs_tiny_x, words_x, roundtrip_bad_x, max_min_rel_x, red_x_nonzero);
    const double near1 = 1.0 - std::ldexp(1.0, -27);
    const uint32_t w1 = frac_word(near1);
    float sr = frac_encode(1.0, 0.0), sb = frac_encode(0.0, 1.0), s0 = frac_encode(0.0, 0.0);
    uint32_t br, bb;
    std::memcpy(&br, &sr, 4);
    std::memcpy(&bb, &sb, 4);
    double zr, zb;
    frac_split(0.0, s0, zr, zb);
    printf("{\"max_rel\": %.6e, \"not_nearest\": %ld, \"max_abs_tiny\": %.6e, \"roundtrip_words\": %ld, "
           "\"roundtrip_bad\": %ld, \"max_min_rel\": %.6e, \"max_maj\": %.6e, \"max_esc\": %.6e, "
           "\"esc_words\": %ld, \"near1_escape\": %d, \"near1_value\": %.17g, \"all_red\": %u, \"all_blue\": %u, "
           "\"red_pure\": %d, \"blue_pure\": %d, \"zero_split\": [%g, %g]}\n",
           max_rel, not_nearest, max_abs_tiny, words, roundtrip_bad, max_min_rel, max_maj, max_esc, esc_words,
           (int)((w1 >> 25) == 127u), frac_value(w1), br, bb, (int)frac_all_red(sr), (int)frac_all_blue(sb), zr, zb);
    return 0;
}
