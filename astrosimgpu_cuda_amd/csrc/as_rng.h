// as_rng.h -- splitmix64 counter RNG shared by host setup and device kernels.
//
// Every random number in a run is a pure function of (seed, stream, counter):
// no generator state lives on the device, so any cell can draw its numbers
// for any step independently and in any order, and the host and the GPU draw
// identical bits. Transcendentals come from as_dmath.h so the floating-point
// draws (normal, Poisson) are also identical on host and device.
#pragma once

#include <cstdint>

#include "as_dmath.h"

#if defined(__HIPCC__)
#define AS_HD __host__ __device__ __forceinline__
#else
#define AS_HD inline
#endif

namespace astrosim {

constexpr std::uint64_t kRngStreamMul = 0x9E3779B97F4A7C15ULL;
constexpr std::uint64_t kRngCounterMul = 0xBF58476D1CE4E5B9ULL;

// The splitmix64 output finaliser.
AS_HD std::uint64_t rng_mix(std::uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// The finaliser's input for counter 0 of a stream; counter c adds
// c * kRngCounterMul (all mod 2^64). Kernels advance these sums instead of
// multiplying per draw: stream s + k*d has base rng_base(seed, s) +
// k*(d*kRngStreamMul).
AS_HD std::uint64_t rng_base(std::uint64_t seed, std::uint64_t stream) {
    return seed + stream * kRngStreamMul;
}

// Reference: src/astrosim_fused.cu:127-135. Weyl increment of the stream and
// counter into the seed, then the splitmix64 output finaliser.
AS_HD std::uint64_t rng_bits(std::uint64_t seed, std::uint64_t stream, std::uint64_t counter) {
    return rng_mix(rng_base(seed, stream) + counter * kRngCounterMul);
}

// A uniform in [0, 1) from 64 random bits: the top 53, exactly scaled.
AS_HD double rng_to_unit(std::uint64_t bits) {
    return static_cast<double>(bits >> 11) * (1.0 / 9007199254740992.0);
}

// Reference: src/astrosim_fused.cu:137-141. Top 53 bits scaled to [0, 1);
// exact in double precision.
AS_HD double rng_uniform(std::uint64_t seed, std::uint64_t stream, std::uint64_t counter) {
    return rng_to_unit(rng_bits(seed, stream, counter));
}

// Box-Muller, cosine branch, on counters 0 and 1 of the stream whose
// counter-0 finaliser input is z (rng_base). Reference:
// src/astrosim_fused.cu:143-149 (u1 clamped at 1e-300).
AS_HD double rng_normal_z(std::uint64_t z) {
    double u1 = rng_to_unit(rng_mix(z));
    if (u1 <= 1e-300) u1 = 1e-300;
    const double u2 = rng_to_unit(rng_mix(z + kRngCounterMul));
    return __builtin_sqrt(-2.0 * as_log(u1)) * as_cos(6.283185307179586476925286766559 * u2);
}

AS_HD double rng_normal(std::uint64_t seed, std::uint64_t stream) {
    return rng_normal_z(rng_base(seed, stream));
}

// Poisson count for a fresh stream (counter-0 finaliser input z). Knuth's
// product-of-uniforms method for lambda <= 30 (counters 0, 1, 2, ...), a
// rounded normal approximation above. `limit` must equal as_exp(-lambda);
// callers hoist it because lambda is a per-population constant.
// Reference: src/astrosim_fused.cu:154-175.
AS_HD int rng_poisson_z(std::uint64_t z, double lambda, double limit) {
    if (lambda <= 0.0) return 0;
    if (lambda > 30.0) {
        const double v = lambda + __builtin_sqrt(lambda) * rng_normal_z(z);
        return v < 0.0 ? 0 : static_cast<int>(v + 0.5);
    }
    // The first test, u0 > limit, is decided in integers: u0 = m0 * 2^-53
    // with integer m0, so u0 > limit <=> m0 > floor(limit * 2^53) (exact
    // scaling; limit is in (0, 1)). Most draws stop here.
    const std::uint64_t m0 = rng_mix(z) >> 11;
    if (m0 <= static_cast<std::uint64_t>(limit * 9007199254740992.0)) return 0;
    double product = static_cast<double>(m0) * (1.0 / 9007199254740992.0);
    int count = 0;
    while (product > limit) {
        ++count;
        z += kRngCounterMul;
        product *= rng_to_unit(rng_mix(z));
    }
    return count;
}

AS_HD int rng_poisson(std::uint64_t seed, std::uint64_t stream, double lambda, double limit) {
    return rng_poisson_z(rng_base(seed, stream), lambda, limit);
}

// Stream-key constants XOR-ed into the run seed, one per draw site.
// Reference: src/astrosim_fused.cu:395 (astro Poisson), :404 (astro noise),
// :476 (neuron Poisson), :499 (neuron noise), :634/:917 (host setup stream).
constexpr std::uint64_t kKeyAstroPoisson = 0xC2B2AE3D27D4EB4FULL;
constexpr std::uint64_t kKeyAstroNoise = 0x9E3779B97F4A7C15ULL;
constexpr std::uint64_t kKeyNeuronPoisson = 0x2545F4914F6CDD1DULL;
constexpr std::uint64_t kKeyNeuronNoise = 0x9E3779B185EBCA87ULL;
constexpr std::uint64_t kSetupStream = 0x5EEDULL;
// Per-step stream ids are step*kStepStride + cell (Poisson and astrocyte
// noise) and noise_idx*kNoiseStride + neuron (neuron noise). They wrap in
// 64 bits and alias across steps for cell >= kStepStride; that is part of
// the draw contract. Reference: src/astrosim_fused.cu:396, :405, :477, :500.
constexpr std::uint64_t kStepStride = 1000003ULL;
constexpr std::uint64_t kNoiseStride = 1000033ULL;

}  // namespace astrosim
