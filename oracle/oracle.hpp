// oracle.hpp — CPU restatement of Helios's path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY. This is the checker the GPU path is compared
// against; nothing in the product library links, loads or calls it. Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// It is a scalar transcription of the reference GLSL programs, templated on
// the real type R: R = float is the FP32 transcription (the "reference path"
// timed as the CPU baseline, and the bit-exact reference of primary-hit IDs);
// R = double is the FP64 converged oracle. Every function cites the shader or
// host file:line it restates (paths relative to the reference root).
//
// Compile with -ffp-contract=off: ray generation, the watertight triangle test
// and the alpha test must round exactly as the GPU kernels do (they are built
// with contraction switched off for the same functions).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/helios_hip.h"

namespace orc {

// ---------------------------------------------------------------------------
// Small vector math in precision R.
// ---------------------------------------------------------------------------
template <typename R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <typename R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <typename R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
// GLSL normalize(v) = v * inversesqrt(dot(v, v)); the GPU kernels use the same
// correctly rounded 1/sqrt so both sides agree.
template <typename R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / std::sqrt(dot(a, a))); }
// GLSL min/max follow IEEE minNum/maxNum on the reference hardware: a NaN
// operand yields the other operand. fmin/fmax have that rule.
template <typename R> inline R gmin(R a, R b) { return std::fmin(a, b); }
template <typename R> inline R gmax(R a, R b) { return std::fmax(a, b); }
template <typename R> inline R gclamp(R x, R lo, R hi) { return gmin(gmax(x, lo), hi); }
template <typename R> inline V3<R> vmin(V3<R> a, R b) { return {gmin(a.x, b), gmin(a.y, b), gmin(a.z, b)}; }
template <typename R> inline R mix(R x, R y, R a) { return x * (R(1) - a) + y * a; }
template <typename R> inline V3<R> mix(V3<R> x, V3<R> y, R a) { return x * (R(1) - a) + y * a; }
template <typename R> inline R smoothstep(R e0, R e1, R x) {
    R t = gclamp((x - e0) / (e1 - e0), R(0), R(1));
    return t * t * (R(3) - R(2) * t);
}
// GLSL reflect(I, N) = I - 2 * dot(N, I) * N
template <typename R> inline V3<R> reflect(V3<R> i, V3<R> n) { return i - R(2) * dot(n, i) * n; }
template <typename R> inline bool is_black(V3<R> c) { return c.x == R(0) && c.y == R(0) && c.z == R(0); }

template <typename R> constexpr R kPi() { return R(3.14159265359); }          // common.glsl:18
template <typename R> constexpr R kEps() { return R(0.0001); }                // common.glsl:19
template <typename R> constexpr R kMinRoughness() { return R(0.1); }          // common.glsl:21

// ---------------------------------------------------------------------------
// RNG: xoroshiro64* + Thomas Wang hash (random.glsl:11-50) and the draw
// helpers (sampling.glsl:6-26). Integer-exact.
// ---------------------------------------------------------------------------
struct RNG {
    uint32_t s0, s1;
};
inline uint32_t rng_rotl(uint32_t x, uint32_t k) { return (x << k) | (x >> (32u - k)); }
inline uint32_t rng_next(RNG& r) {
    uint32_t result = r.s0 * 0x9e3779bbu;
    r.s1 ^= r.s0;
    r.s0 = rng_rotl(r.s0, 26) ^ r.s1 ^ (r.s1 << 9);
    r.s1 = rng_rotl(r.s1, 13);
    return result;
}
inline uint32_t rng_hash(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}
inline RNG rng_init(uint32_t x, uint32_t y, uint32_t frame) {
    RNG r;
    r.s0 = rng_hash((x << 16) | y);
    r.s1 = rng_hash(frame);
    rng_next(r);
    return r;
}
inline float next_float(RNG& r) {
    uint32_t u = 0x3f800000u | (rng_next(r) >> 9);
    float f;
    std::memcpy(&f, &u, 4);
    return f - 1.0f;
}
inline uint32_t next_uint(RNG& r, uint32_t nmax) {
    float f = next_float(r);
    return (uint32_t)std::floor(f * (float)nmax);
}

// ---------------------------------------------------------------------------
// Scene, as packed by Scene::create_gpu_resources (scene.cpp:915-1311).
// ---------------------------------------------------------------------------
struct MaterialData {  // scene.cpp:25-32
    int32_t tex0[4];   // albedo, normals, roughness, metallic
    int32_t tex1[4];   // emissive, -, roughness channel, metallic channel
    float albedo[4];
    float emissive[4];
    float roughness_metallic[4];
};
struct LightData {     // scene.cpp:36-42
    float d0[4], d1[4], d2[4], d3[4];
};
struct InstanceData {  // scene.cpp:46-52 (+ the world->object 3x4 used by traversal)
    float model[16];
    float normal[16];
    uint32_t mesh_index;
    float w2o[12];     // row-major 3x4 inverse of model
};
struct Texture {
    uint32_t w = 0, h = 0, format = 0;
    std::vector<float> texels;  // RGBA float, decoded (sRGB -> linear for rgb)
};
struct Mesh {
    std::vector<hh_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<hh_submesh> submeshes;
    std::vector<uint32_t> materials;  // material ids
    // BLAS (CPU BVH2 over the triangles of all submeshes)
    struct Tri { float v[3][3]; uint32_t geom, prim; };
    std::vector<Tri> tris;
    struct Node { float lo[3], hi[3]; int32_t left, right; uint32_t start, count; };
    std::vector<Node> nodes;
    std::vector<uint8_t> geom_opaque;  // mesh.cpp:75-76
};

struct Scene {
    // user inputs
    std::vector<Texture> textures;
    std::vector<hh_material_desc> materials;
    std::vector<Mesh> meshes;
    std::vector<hh_instance_desc> instances;
    std::vector<hh_light_desc> lights;
    std::vector<float> ibl;  // 6 * size * size * 4
    uint32_t ibl_size = 0;
    hh_sky_params sky{4.0f, 0.1f, 1.15f, 512, 0};
    // packed
    std::vector<MaterialData> mat_data;
    std::vector<int32_t> tex_slots;          // descriptor slot -> texture id
    std::vector<LightData> light_data;
    std::vector<InstanceData> inst_data;
    std::vector<std::vector<uint32_t>> submesh_info;  // per instance: (prim_offset, mat) pairs
    std::vector<float> inst_world_lo, inst_world_hi;  // per instance world AABB
    uint32_t num_lights = 0;
    // environment cube (float texels, 6 faces), size env_size
    std::vector<float> env;
    uint32_t env_size = 0;
    // extension: equirectangular RGBA16F sky (bits), width env_eq_w (0 = cube)
    std::vector<uint16_t> env_eq;
    uint32_t env_eq_w = 0;
    // extension: power-proportional emitters (every emissive triangle of
    // every instance): world vertices, area, emission, probability, CDF
    std::vector<float> emit_v;     // 9 per triangle
    std::vector<float> emit_area, emit_p, emit_cdf;
    std::vector<float> emit_le;    // 3 per triangle
    uint32_t num_area = 0;         // area-light records at the front of light_data
    float sky_coef[30] = {0};
    bool committed = false;
};

// Host-side helpers implemented in oracle.cpp
void commit(Scene& s);
void launch_consts(const hh_camera& cam, uint32_t w, uint32_t h, uint32_t num_lights, uint32_t frame,
                   const hh_settings& st, hh_launch_consts& out);

}  // namespace orc
