// Colour split of an FP32-stored population f = fr + fb: one 32-bit word
// holding the fraction r of its minority colour (the colour of smaller
// magnitude, r = minority / f) and, in bit 0, which colour that is (1: blue).
//
// For two parts of one sign r lies in [0, 1/2], so a float's sign bit and
// its exponent's top bit would always be zero; the word spends them on two
// more mantissa bits instead:
//   bits 31..25  e: r's float exponent field (0..126)
//   bits 24..1   m: 24 mantissa bits
//   bit  0       colour of the minority part
//   r = (2^24 + m) 2^(e - 151)   for e >= 1,   r = m 2^-150   for e = 0,
// i.e. r rounded to nearest with a 25-bit significand (relative error at
// most 2^-25 down to r = 2^-126, absolute at most 2^-151 below).  Parts of
// opposite sign, or a fraction that rounds to 1, take the escape e = 127:
// bit 24 sign, bits 23..16 the float exponent field, bits 15..1 the top 15
// mantissa bits of r (relative error at most 2^-16; a fraction beyond the
// float range reads back as infinite).
// Single-colour populations are exact: 0x00000001 is all red (blue part 0),
// 0x00000000 all blue.
//
// The extended form (XB = 8) adds one byte x per population with the next 8
// significand bits: 33 bits in all (relative error at most 2^-33 down to
// 2^-126, absolute 2^-159 below); escape words keep x = 0.  The word of the
// extended form is r rounded at 33 bits and cut to its first 25.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace psx_lbm_dev {

__host__ __device__ inline uint64_t frac_dbits(double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    return b;
}
// 2^e for -1022 <= e <= 1023
__host__ __device__ inline double frac_pow2(int e) {
    const uint64_t b = (uint64_t)(e + 1023) << 52;
    double x;
    std::memcpy(&x, &b, 8);
    return x;
}

// The word of fraction r (bit 0 clear) and, for XB = 8, its extension byte
// x; branch-free.  On [2^-126, 1) the double's own bits are rounded at the
// (25 + XB)-th significand bit (half-way cases away from zero; the carry
// runs into the exponent) and repacked, below 2^-126 r is rounded to a
// multiple of 2^-(150 + XB), else the escape is formed.
template <int XB>
__host__ __device__ inline uint32_t frac_word_t(double r, uint32_t &x) {
    static_assert(XB == 0 || XB == 8, "XB is 0 or 8");
    const uint64_t rb = frac_dbits(r) + (1ull << (27 - XB));
    const uint32_t hi = (uint32_t)(rb >> 32), lo = (uint32_t)rb;
    const uint32_t wn = (hi << 5) - (896u << 25) + ((lo >> 27) & 0x1Eu);
    const bool normal = hi - (897u << 20) < (126u << 20);  // 2^-126 <= r (rounded) < 1
    const bool small = hi < (897u << 20);                  // 0 <= r < 2^-126
    // r 2^(150+XB) < 2^32 wherever small holds; the clamp keeps the
    // conversion defined elsewhere (the value is then unused)
    const uint32_t n = (uint32_t)std::fmin(std::fmax(std::rint(r * frac_pow2(150 + XB)), 0.0), 4294967295.0);
    const uint32_t wd = (n >> XB) << 1;
    const float s = (float)r;
    uint32_t b;
    std::memcpy(&b, &s, 4);
    b += 0x7Fu + ((b >> 8) & 1u);  // round to nearest even at 15 mantissa bits
    const uint32_t wx = (127u << 25) | ((b >> 7) & ~1u);
    if (XB) {
        const uint32_t xn = (lo >> (28 - XB)) & ((1u << XB) - 1u);
        const uint32_t xd = n & ((1u << XB) - 1u);
        x = normal ? xn : (small ? xd : 0u);
    } else {
        x = 0u;
    }
    return normal ? wn : (small ? wd : wx);
}
__host__ __device__ inline uint32_t frac_word(double r) {
    uint32_t x;
    return frac_word_t<0>(r, x);
}

// The fraction held by a word (and extension byte x) that is not an escape
// word: the double (1 + m 2^-24 + x 2^-32) 2^(e-127) assembled from the
// bits, and for e = 0 2 (that - 2^-127).  Sets esc when w is an escape word
// (whose value this is not).
__host__ __device__ inline double frac_value_fast(uint32_t w, uint32_t x, bool &esc) {
    const uint64_t hi = (uint64_t)((w >> 5) + (896u << 20));
    const uint64_t lo = (uint64_t)(((w << 27) & 0xF0000000u) | (x << 20));
    double d;
    const uint64_t bits = (hi << 32) | lo;
    std::memcpy(&d, &bits, 8);
    esc |= (w >> 25) == 127u;
    return (w >> 25) != 0u ? d : std::fma(d, 2.0, -frac_pow2(-126));
}

// the fraction held by any word (and extension byte)
__host__ __device__ inline double frac_value(uint32_t w, uint32_t x = 0u) {
    bool esc = false;
    const double v = frac_value_fast(w, x, esc);
    const uint32_t b = (((w >> 24) & 1u) << 31) | (((w >> 16) & 0xFFu) << 23) | (((w >> 1) & 0x7FFFu) << 8);
    float s;
    std::memcpy(&s, &b, 4);
    return esc ? (double)s : v;
}

// stored word of the population (fr, fb), as a float's bit pattern, and
// for XB = 8 its extension byte
template <int XB>
__host__ __device__ inline float frac_encode_t(double fr, double fb, uint32_t &x) {
    const double f = fr + fb;
    const bool blue = std::fabs(fb) < std::fabs(fr);
    // correctly rounded division: an approximate reciprocal biases the split
    const double r = f == 0.0 ? 0.5 : (blue ? fb : fr) / f;
    const uint32_t w = frac_word_t<XB>(r, x) | (blue ? 1u : 0u);
    float out;
    std::memcpy(&out, &w, 4);
    return out;
}
__host__ __device__ inline float frac_encode(double fr, double fb) {
    uint32_t x;
    return frac_encode_t<0>(fr, fb, x);
}

// red and blue parts of the population f whose stored word is s (and
// extension byte x)
__host__ __device__ inline void frac_split(double f, float s, double &fr, double &fb, uint32_t x = 0u) {
    uint32_t w;
    std::memcpy(&w, &s, 4);
    const double m = frac_value(w, x) * f, o = f - m;
    const bool blue = w & 1u;
    fr = blue ? o : m;
    fb = blue ? m : o;
}
// the same without the escape decode: exact unless esc comes back set
__host__ __device__ inline void frac_split_fast(double f, float s, uint32_t x, double &fr, double &fb, bool &esc) {
    uint32_t w;
    std::memcpy(&w, &s, 4);
    const double m = frac_value_fast(w, x, esc) * f, o = f - m;
    const bool blue = w & 1u;
    fr = blue ? o : m;
    fb = blue ? m : o;
}

// the population is all red / all blue
__host__ __device__ inline bool frac_all_red(float s) {
    uint32_t w;
    std::memcpy(&w, &s, 4);
    return w == 1u;
}
__host__ __device__ inline bool frac_all_blue(float s) {
    uint32_t w;
    std::memcpy(&w, &s, 4);
    return w == 0u;
}

}  // namespace psx_lbm_dev
