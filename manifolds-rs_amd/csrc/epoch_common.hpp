// Device/host helpers for the Adam layout epoch (shared by the epoch kernels
// and the multi-fit driver).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrs {

// Constants of the "adam_gpu" gradient, src/training/umap_optimiser_gpu.rs:58-70.
constexpr float kGradClipVal = 4.0f;
constexpr float kGradRepEps = 0.001f;
constexpr float kGradDistSqThreshold = 1e-8f;
// DENS_EPS, src/utils/density.rs:38
constexpr double kDensEps = 1e-8;

// Exact u32 remainder by a run-time invariant divisor d (1 <= d < 2^32):
// with M = ceil(2^64 / d), a mod d = ((M * a mod 2^64) * d) >> 64 for every
// a < 2^32 (Lemire, Kaser & Kurz, "Faster Remainder by Direct Computation",
// Software: Practice and Experience 49(6), 2019, Theorem 1).  Replaces the
// ~30-instruction integer-remainder expansion in the negative-sample hash.
struct FastMod {
    uint64_t m;
    uint32_t d;
};

__host__ __device__ inline FastMod make_fastmod(uint32_t d) {
    FastMod f;
    f.d = d;
    f.m = UINT64_C(0xFFFFFFFFFFFFFFFF) / d + 1;
    return f;
}

__host__ __device__ inline uint32_t fastmod(uint32_t a, FastMod f) {
    uint64_t low = f.m * (uint64_t)a;
    uint32_t lo = (uint32_t)low, hi = (uint32_t)(low >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t lo_part = __umulhi(lo, f.d);
#else
    uint32_t lo_part = (uint32_t)(((uint64_t)lo * f.d) >> 32);
#endif
    return (uint32_t)(((uint64_t)hi * f.d + lo_part) >> 32);
}

// Splitmix-style negative-sample hash of the reference, bit for bit:
// gpu_hash_neg, src/training/umap_optimiser_gpu.rs:611-624.
__host__ __device__ inline uint32_t hash_neg(uint32_t seed, uint32_t node, uint32_t epoch,
                                             uint32_t edge_local, uint32_t neg, FastMod n) {
    uint32_t h = seed ^ (node * 0x9E3779B1u) ^ (epoch * 0x85EBCA77u) ^
                 (edge_local * 0xC2B2AE3Du) ^ (neg * 0x27D4EB2Fu);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return fastmod(h, n);
}

// Per-epoch host scalars of launch_adam_update (umap_optimiser_gpu.rs:1168-1175),
// computed in f32 exactly as the reference does.
struct AdamScalars {
    float fused_lr;  // lr * (1 - e/N) * sqrt(1 - beta2^t) / (1 - beta1^t)
    float epsc;      // sqrt(1 - beta2^t) * eps
};

}  // namespace mrs
