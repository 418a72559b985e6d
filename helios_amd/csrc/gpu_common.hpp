// gpu_common.hpp — device math, RNG, sampling and the uber BRDF for gfx950.
//
// Restates, for the GPU, the reference shader library:
//   random.glsl:11-50, sampling.glsl:6-36, brdf.glsl:6-167, common.glsl:130-141.
// The library is compiled with -ffp-contract=off and correctly rounded fp32
// division/sqrt, so the functions on the exact-parity path (RNG, ray
// generation, watertight intersection, alpha test) round exactly like their
// CPU restatement; hot arithmetic elsewhere uses explicit fmaf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HH_PI 3.14159265359f
#define HH_EPS 0.0001f
#define HH_MIN_ROUGHNESS 0.1f

namespace hh {

// ---------------------------------------------------------------------------
// float3 helpers
// ---------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__host__ __device__ inline float length(f3 a) { return sqrtf(dot(a, a)); }
__host__ __device__ inline f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__host__ __device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__host__ __device__ inline bool is_black(f3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
__host__ __device__ inline f3 reflect(f3 i, f3 n) { return i - 2.0f * dot(n, i) * n; }
__device__ inline f3 ld3(const float4 v) { return f3{v.x, v.y, v.z}; }
__device__ inline f3 ld3p(const float* p) { return f3{p[0], p[1], p[2]}; }

// ---------------------------------------------------------------------------
// RNG (random.glsl:11-50, sampling.glsl:6-26)
// ---------------------------------------------------------------------------
struct RNG {
    uint32_t s0, s1;
};
__host__ __device__ inline uint32_t rng_rotl(uint32_t x, uint32_t k) { return (x << k) | (x >> (32u - k)); }
__host__ __device__ inline uint32_t rng_next(RNG& r) {
    uint32_t result = r.s0 * 0x9e3779bbu;
    r.s1 ^= r.s0;
    r.s0 = rng_rotl(r.s0, 26) ^ r.s1 ^ (r.s1 << 9);
    r.s1 = rng_rotl(r.s1, 13);
    return result;
}
__host__ __device__ inline uint32_t rng_hash(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}
__host__ __device__ inline RNG rng_init(uint32_t x, uint32_t y, uint32_t frame) {
    RNG r;
    r.s0 = rng_hash((x << 16) | y);
    r.s1 = rng_hash(frame);
    rng_next(r);
    return r;
}
__host__ __device__ inline float next_float(RNG& r) {
    uint32_t u = 0x3f800000u | (rng_next(r) >> 9);
    return __uint_as_float(u) - 1.0f;
}
__host__ __device__ inline uint32_t next_uint(RNG& r, uint32_t nmax) {
    float f = next_float(r);
    return (uint32_t)floorf(f * (float)nmax);
}

// ---------------------------------------------------------------------------
// BRDF set (brdf.glsl)
// ---------------------------------------------------------------------------
struct Surface {
    f3 pos, vnormal, normal, albedo, emissive, F0;
    float alpha_a;  // albedo.a
    float metallic, roughness;
};

// sampling.glsl:28-36 — make_rotation_matrix(z) * t
__device__ inline f3 rot_mul(f3 z, f3 t) {
    f3 ref = fabsf(dot(z, mk(0.f, 1.f, 0.f))) > 0.99f ? mk(0.f, 0.f, 1.f) : mk(0.f, 1.f, 0.f);
    f3 x = normalize(cross(ref, z));
    f3 y = cross(z, x);
    return x * t.x + y * t.y + z * t.z;
}
__device__ inline f3 sample_cosine_lobe(f3 n, float r0, float r1) {  // brdf.glsl:6-18
    float rx = fmaxf(0.00001f, r0), ry = fmaxf(0.00001f, r1);
    float phi = 2.0f * HH_PI * ry;
    float ct = sqrtf(rx), st = sqrtf(1.0f - rx);
    return normalize(rot_mul(n, mk(st * cosf(phi), st * sinf(phi), ct)));
}
__device__ inline float D_ggx(float ndoth, float alpha) {  // brdf.glsl:58-64
    float a2 = alpha * alpha;
    float denom = (ndoth * ndoth) * (a2 - 1.0f) + 1.0f;
    return a2 / fmaxf(HH_EPS, HH_PI * denom * denom);
}
__device__ inline float G1_schlick_ggx(float roughness, float ndotv) {  // brdf.glsl:66-71
    float k = ((roughness + 1.0f) * (roughness + 1.0f)) / 8.0f;
    return ndotv / fmaxf(HH_EPS, ndotv * (1.0f - k) + k);
}
__device__ inline f3 F_schlick(f3 f0, float vdoth) {
    float w = powf(1.0f - vdoth, 5.0f);
    return f0 + (mk(1.f, 1.f, 1.f) - f0) * w;
}
__device__ inline f3 sample_ggx(f3 n, float alpha, float x0, float x1) {  // brdf.glsl:83-96
    float phi = 2.0f * HH_PI * x0;
    float ct = sqrtf((1.0f - x1) / (1.0f + (alpha * alpha - 1.0f) * x1));
    float st = sqrtf(1.0f - ct * ct);
    return normalize(rot_mul(n, mk(st * cosf(phi), st * sinf(phi), ct)));
}
__device__ inline f3 evaluate_uber(const Surface& p, f3 Wo, f3 Wh, f3 Wi) {  // brdf.glsl:109-121
    float NdotL = fmaxf(dot(p.normal, Wi), 0.0f);
    float NdotV = fmaxf(dot(p.normal, Wo), 0.0f);
    float NdotH = fmaxf(dot(p.normal, Wh), 0.0f);
    float VdotH = fmaxf(dot(Wi, Wh), 0.0f);
    f3 F = F_schlick(p.F0, VdotH);
    float a = p.roughness * p.roughness;
    float G = G1_schlick_ggx(p.roughness, NdotL) * G1_schlick_ggx(p.roughness, NdotV);
    f3 spec = ((F * D_ggx(NdotH, a)) * G) / fmaxf(HH_EPS, 4.0f * NdotL * NdotV);
    f3 diffuse = p.albedo / HH_PI;
    return (mk(1.f, 1.f, 1.f) - F) * diffuse + spec;
}
__device__ inline float pdf_uber(const Surface& p, f3 Wh, f3 Wi) {  // brdf.glsl:123-134
    float NdotL = fmaxf(dot(p.normal, Wi), 0.0f);
    float NdotH = fmaxf(dot(p.normal, Wh), 0.0f);
    float VdotH = fmaxf(dot(Wi, Wh), 0.0f);
    float pd = NdotL / HH_PI;
    float ps = D_ggx(NdotH, p.roughness * p.roughness) * NdotH / fmaxf(HH_EPS, 4.0f * VdotH);
    return pd * (1.0f - 0.5f) + ps * 0.5f;
}
// brdf.glsl:136-167 — the RNG is passed by value (reference quirk: the lobe
// draws do not advance the path's stream).
__device__ inline f3 sample_uber(const Surface& p, f3 Wo, RNG rng, f3& Wi, float& pdf) {
    float alpha = p.roughness * p.roughness;
    float rx = next_float(rng);
    float ry = next_float(rng);
    float rz = next_float(rng);
    f3 Wh;
    bool spec = false;
    if (rx < 0.5f) {
        Wh = sample_ggx(p.normal, alpha, ry, rz);
        Wi = reflect(-Wo, Wh);
        float NdotL = fmaxf(dot(p.normal, Wi), 0.0f);
        float NdotV = fmaxf(dot(p.normal, Wo), 0.0f);
        if (NdotL > 0.0f && NdotV > 0.0f) spec = true;
    }
    if (!spec) {
        Wi = sample_cosine_lobe(p.normal, ry, rz);
        Wh = normalize(Wo + Wi);
    }
    pdf = pdf_uber(p, Wh, Wi);
    return evaluate_uber(p, Wo, Wh, Wi);
}

}  // namespace hh
