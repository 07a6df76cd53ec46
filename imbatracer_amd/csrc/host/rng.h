// MWC64X generator and the seed scramble of the reference's ray generators.
//   RNG                 render/random.h:15-47
//   seed scramble       render/ray_gen/ray_gen.h:65-72 (pixels), 139-145 (light paths)
// The device copy lives in hip/rng_dev.h; both must produce the same stream.
#pragma once

#include <cstdint>

namespace imba {

struct Mwc64x {
    uint64_t state;
    explicit Mwc64x(uint64_t seed = 0) : state(seed) {}
    uint32_t next() {
        const uint32_t c = (uint32_t)(state >> 32);
        const uint32_t x = (uint32_t)(state & 0xFFFFFFFFu);
        state = (uint64_t)x * 4294883355ull + c;
        return x ^ c;
    }
    // x / 0xFFFFFFFF in float: [0, 1], both ends possible after rounding.
    float uniform() { return (float)next() / (float)0xFFFFFFFFu; }
    // lerp(lo, hi, r) = lo * (1 - r) + hi * r (core/common.h:110-113)
    float uniform(float lo, float hi) {
        const float r = uniform();
        return lo * (1.0f - r) + hi * r;
    }
    int uniform_int(int lo, int hi) { return hi == lo ? lo : (int)(next() % (uint32_t)(hi - lo)) + lo; }
    void discard(int n) {
        for (int i = 0; i < n; i++) next();
    }
};

// Bernstein-style scramble of a base seed with a sample index, then the discard rule.
inline Mwc64x seeded_stream(uint32_t seed_base, int index, int pixel_or_path) {
    int seed = (int)seed_base;
    for (int k = 0; k < 4; k++) seed = (int)((uint32_t)(33u * (uint32_t)seed) ^ (uint32_t)index);
    Mwc64x r((uint64_t)(int64_t)seed);
    r.discard((seed % 5) + 16 + pixel_or_path % 5);
    return r;
}

}  // namespace imba
