// synth.h -- seeded synthetic RGBA8 textures, integer-only so that the host
// and the device produce the same bytes.
//
// Kind 0 ("waves"): per channel, base 128 plus three 2-D sinusoids with
//   periods log-uniform in [8, 256] px and random direction and phase, plus
//   uniform integer noise in [-noise, +noise], saturated to [0, 255];
//   alpha = 255.  The sine is a parabolic approximation evaluated on a 32-bit
//   phase (2^32 = one period), so every pixel is a pure integer function of
//   (parameters, x, y).
// Kind 1 ("stress"): 64x64 regions, each drawn from {constant colour,
//   two-colour checker, saturated primaries at 0/255, uniform u8 noise,
//   normal-map field (unit xyz vectors mapped to [0,255])}.
// Kind 2 ("waves + alpha"): the colour of kind 0 and an alpha channel made of
//   64x64 regions, each opaque (255), a binary cutout (0 or 255 where a
//   three-sinusoid field crosses its midpoint) or that smooth field itself
//   plus the same uniform noise, in proportions 1 : 1 : 2.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define GOOFY_HD __host__ __device__ inline
#else
#define GOOFY_HD inline
#endif

namespace goofy_synth {

struct Params {
    uint32_t kind;        // 0 waves, 1 stress
    uint32_t seed;        // per-texture seed (mixed from the user seed and texture id)
    uint32_t noise;       // waves: noise amplitude
    int32_t base[3];      // waves: per-channel base level
    int32_t amp[3][3];    // waves: amplitudes (pixel units)
    uint32_t phase[3][3]; // waves: phase offsets
    uint32_t kx[3][3];    // waves: phase increment per pixel along x (wrapping)
    uint32_t ky[3][3];    // waves: phase increment per pixel along y (wrapping)
    int32_t aamp[3];      // alpha field (kind 2): amplitudes, phases, increments
    uint32_t aphase[3], akx[3], aky[3];
};

// 32-bit integer hash (xorshift-multiply finaliser).
GOOFY_HD uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

GOOFY_HD uint32_t hash3(uint32_t seed, uint32_t a, uint32_t b, uint32_t c)
{
    return mix32(seed ^ mix32(a + 0x9e3779b9u * mix32(b + 0x85ebca6bu * mix32(c + 0x68e31da4u))));
}

// Parabolic sine on a 32-bit phase: result in [-16384, 16384] (Q14).
GOOFY_HD int32_t isin(uint32_t phase)
{
    const int32_t x = (int32_t)(phase >> 16) - ((phase & 0x80000000u) ? 65536 : 0);  // [-32768, 32767]
    const int32_t ax = x < 0 ? -x : x;
    return (x * (32768 - ax)) / 16384;
}

GOOFY_HD uint32_t isqrt32(uint32_t v)
{
    uint32_t r = 0, bit = 1u << 30;
    while (bit > v) bit >>= 2;
    while (bit) {
        if (v >= r + bit) {
            v -= r + bit;
            r = (r >> 1) + bit;
        } else {
            r >>= 1;
        }
        bit >>= 2;
    }
    return r;
}

GOOFY_HD uint32_t clamp255(int32_t v) { return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

GOOFY_HD uint32_t pixel_waves(const Params& P, uint32_t x, uint32_t y)
{
    uint32_t out = 0xFF000000u;
    for (int c = 0; c < 3; ++c) {
        int32_t v = P.base[c];
        for (int k = 0; k < 3; ++k) {
            const uint32_t ph = P.phase[c][k] + P.kx[c][k] * x + P.ky[c][k] * y;
            v += (P.amp[c][k] * isin(ph)) / 16384;
        }
        if (P.noise) {
            const uint32_t h = hash3(P.seed, x, y, (uint32_t)c);
            v += (int32_t)(((h >> 16) * (2u * P.noise + 1u)) >> 16) - (int32_t)P.noise;
        }
        out |= clamp255(v) << (8 * c);
    }
    return out;
}

GOOFY_HD uint32_t pixel_stress(const Params& P, uint32_t x, uint32_t y)
{
    const uint32_t rx = x >> 6, ry = y >> 6;
    const uint32_t h = hash3(P.seed, rx, ry, 0xA11CEu);
    const uint32_t kind = ((h >> 8) & 0xFFFFu) * 5u >> 16;  // 0..4
    switch (kind) {
    case 0:  // constant colour
        return 0xFF000000u | (mix32(h ^ 1u) & 0x00FFFFFFu);
    case 1: {  // two-colour checker, cell 1, 2 or 4 px
        const uint32_t s = h % 3u;
        const uint32_t a = mix32(h ^ 2u) & 0x00FFFFFFu, b = mix32(h ^ 3u) & 0x00FFFFFFu;
        return 0xFF000000u | ((((x >> s) + (y >> s)) & 1u) ? b : a);
    }
    case 2: {  // saturated primaries: every channel 0 or 255
        const uint32_t q = hash3(P.seed, x, y, 7u);
        return 0xFF000000u | ((q & 1u) ? 0xFFu : 0u) | ((q & 2u) ? 0xFF00u : 0u) | ((q & 4u) ? 0xFF0000u : 0u);
    }
    case 3:  // uniform u8 noise
        return 0xFF000000u | (hash3(P.seed, x, y, 9u) & 0x00FFFFFFu);
    default: {  // normal map: n = (nx, ny, sqrt(1 - nx^2 - ny^2)), |nx|,|ny| <= 0.7
        const uint32_t k1 = 0x01000000u + (mix32(h ^ 4u) & 0x03FFFFFFu);  // periods ~16..64 px
        const uint32_t k2 = 0x01000000u + (mix32(h ^ 5u) & 0x03FFFFFFu);
        const int32_t nx = (11469 * isin(mix32(h ^ 6u) + k1 * x + (k2 >> 1) * y)) / 16384;
        const int32_t ny = (11469 * isin(mix32(h ^ 7u) + (k1 >> 2) * x + k2 * y)) / 16384;
        const uint32_t rest = (1u << 28) - (uint32_t)(nx * nx) - (uint32_t)(ny * ny);
        const uint32_t nz = isqrt32(rest);  // <= 16384
        // n in [-16384, 16384] -> round(255 * (n + 1) / 2): always in [0, 255],
        // so no clamp.  (A clamp-and-pack of two channels here was turned by
        // the gfx950 backend into v_ashr_pk_u8_i32 followed by an add that
        // assumed the instruction's upper 16 result bits were zero; on the
        // device they were not.  Unsigned, clamp-free arithmetic avoids it.)
        const uint32_t r = ((uint32_t)(nx + 16384) * 255u + 16384u) >> 15;
        const uint32_t g = ((uint32_t)(ny + 16384) * 255u + 16384u) >> 15;
        const uint32_t b = ((nz + 16384u) * 255u + 16384u) >> 15;
        return 0xFF000000u | r | (g << 8) | (b << 16);
    }
    }
}

GOOFY_HD uint32_t pixel_alpha(const Params& P, uint32_t x, uint32_t y)
{
    const uint32_t cls = (hash3(P.seed, x >> 6, y >> 6, 0xA1FAu) >> 8) & 3u;
    if (cls == 0) return 255u;
    int32_t v = 128;
    for (int k = 0; k < 3; ++k) v += (P.aamp[k] * isin(P.aphase[k] + P.akx[k] * x + P.aky[k] * y)) / 16384;
    if (cls == 1) return v >= 128 ? 255u : 0u;
    if (P.noise) {
        const uint32_t h = hash3(P.seed, x, y, 3u);
        v += (int32_t)(((h >> 16) * (2u * P.noise + 1u)) >> 16) - (int32_t)P.noise;
    }
    return clamp255(v);
}

GOOFY_HD uint32_t pixel(const Params& P, uint32_t x, uint32_t y)
{
    if (P.kind == 2) return (pixel_waves(P, x, y) & 0x00FFFFFFu) | (pixel_alpha(P, x, y) << 24);
    return P.kind == 0 ? pixel_waves(P, x, y) : pixel_stress(P, x, y);
}

// splitmix64 stream for parameter draws (host side).
struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
};

// Parameters for texture `tex` of a seeded set.  Periods: P = 8 * 2^(5u),
// u uniform, with 2^-f approximated by 1 - f/2 on each octave.
inline Params make_params(uint32_t kind, uint64_t seed, uint32_t tex, uint32_t noise)
{
    Params P{};
    SplitMix sm{seed * 0x100000001b3ull ^ ((uint64_t)tex << 17) ^ 0x5EEDull};
    P.kind = kind;
    P.seed = (uint32_t)sm.next();
    P.noise = noise;
    for (int c = 0; c < 3; ++c) {
        P.base[c] = 128;
        for (int k = 0; k < 3; ++k) {
            P.amp[c][k] = 12 + (int32_t)(sm.next() % 29u);  // 12..40 each
            P.phase[c][k] = (uint32_t)sm.next();
            const uint32_t u = (uint32_t)(sm.next() & 0xFFFFu);
            const uint32_t t = 5u * u;  // Q16 in [0, 5)
            const uint32_t oct = t >> 16, f = t & 0xFFFFu;
            const int64_t M = (int64_t)(1u << 29) >> oct;  // 2^32 / 8 / 2^oct
            const int64_t mag = M - ((M >> 1) * (int64_t)f >> 16);
            const uint32_t theta = (uint32_t)sm.next();
            const int64_t cs = isin(theta + 0x40000000u), sn = isin(theta);
            P.kx[c][k] = (uint32_t)(int32_t)((mag * cs) / 16384);
            P.ky[c][k] = (uint32_t)(int32_t)((mag * sn) / 16384);
        }
    }
    // Alpha field, drawn after the colour so that kinds 0 and 1 are unchanged.
    for (int k = 0; k < 3; ++k) {
        P.aamp[k] = 20 + (int32_t)(sm.next() % 41u);  // 20..60 each
        P.aphase[k] = (uint32_t)sm.next();
        const uint32_t t = 5u * (uint32_t)(sm.next() & 0xFFFFu);
        const int64_t M = (int64_t)(1u << 29) >> (t >> 16);
        const int64_t mag = M - ((M >> 1) * (int64_t)(t & 0xFFFFu) >> 16);
        const uint32_t theta = (uint32_t)sm.next();
        P.akx[k] = (uint32_t)(int32_t)((mag * isin(theta + 0x40000000u)) / 16384);
        P.aky[k] = (uint32_t)(int32_t)((mag * isin(theta)) / 16384);
    }
    return P;
}

}  // namespace goofy_synth
