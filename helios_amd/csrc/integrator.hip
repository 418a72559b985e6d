// integrator.hip — the path integrator as wavefront kernels for gfx950.
//
// Replaces the Vulkan ray-tracing pipeline of Helios (raygen
// path_trace_rgen.glsl:132-250, closest-hit path_trace_rchit.glsl:136-580,
// any-hit path_trace_rahit.glsl:174-188, miss path_trace_rmiss.glsl:38-67,
// shadow hit/miss path_trace_shadow.{rchit,rmiss}:16-19) launched once per
// frame by PathIntegrator::launch_rays (src/engine/gfx/path_integrator.cpp:199).
//
// Per frame, for every path slot of every view of a batch:
//   generate   : RNG init + camera ray (4 draws)        -> extension queue 0
//   for depth d: extend  (closest hit, any-hit alpha only at d = 0)
//                shade   (emission, NEE light sample, uber BRDF bounce, RR)
//                        -> shadow queue, extension queue d+1 (compacted with
//                           wave ballots: one atomic per wave)
//                shadow  (visibility; adds the NEE contribution if unoccluded)
//   accumulate : min(L, 1) + progressive mean into the RGBA32F surface.
// Every kernel is persistent (grid = CUs x blocks_per_cu, grid-stride over the
// queue length read from device memory), so the launch sequence of a frame is
// fixed and is captured once as a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gpu_common.hpp"
#include "gpu_scene.hpp"
#include "internal.hpp"
#include "traverse.hpp"

namespace hh {

// ---------------------------------------------------------------------------
// Shared device helpers
// ---------------------------------------------------------------------------
__device__ inline uint32_t lane_id() { return __lane_id(); }

// Wave-aggregated allocation that orders the wave's entries by a small key
// (lanes with a smaller key first, then by lane): one atomic per wave plus
// NK ballots. Used to place rays of the same direction octant next to each
// other, so the lanes a traversal wave refills at once start coherent.
template <uint32_t NK>
__device__ inline uint32_t wave_alloc_keyed(uint32_t* counter, bool want, uint32_t key) {
    const unsigned long long m = __ballot(want);
    if (m == 0ull) return 0u;
    const uint32_t lane = lane_id();
    const unsigned long long below = lane == 0 ? 0ull : ((1ull << lane) - 1ull);
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t k = 0; k < NK; k++) {
        const unsigned long long mk = __ballot(want && key == k);
        if (k < key) pos += (uint32_t)__popcll(mk);
        else if (k == key) pos += (uint32_t)__popcll(mk & below);
    }
    const uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    return base + pos;
}
__device__ inline uint32_t octant(f3 d) { return (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u); }

// Queue segment a wave allocates into and starts reading from (blocks are
// whole waves).
__device__ inline uint32_t wave_shard() { return ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % kQShards; }
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;

__device__ inline uint32_t find_view(const ViewGPU* views, uint32_t nv, uint32_t slot) {
    uint32_t lo = 0, hi = nv - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (views[mid].slot_base <= slot) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct SlotInfo {
    uint32_t view, x, y, sample;
};
// Path slots of a view are numbered pixel-major: the spp samples of one pixel
// are consecutive, so the rays of a wave (64 consecutive slots of the first
// queue) start from the same or neighbouring pixels and traverse the same
// nodes. The pixels follow in row-major order, except at one sample per
// pixel on a tile of whole 8x8 blocks, where they follow in 8x8 blocks (in
// row-major block order), so that a wave's 64 rays leave a square of pixels
// instead of a row segment. Results do not depend on the numbering.
__device__ inline bool block_order(const ViewGPU& v) { return v.spp == 1u && ((v.tile_w | v.tile_h) & 7u) == 0u; }
// row-major tile pixel p -> its place in the slot order
__device__ inline uint32_t pixel_order(const ViewGPU& v, uint32_t p) {
    if (!block_order(v)) return p;
    const uint32_t y = p / v.tile_w, x = p - y * v.tile_w;
    return ((y >> 3) * (v.tile_w >> 3) + (x >> 3)) * 64u + (y & 7u) * 8u + (x & 7u);
}
__device__ inline SlotInfo slot_info(const ViewGPU* views, uint32_t nv, uint32_t slot) {
    SlotInfo si;
    si.view = nv == 1 ? 0u : find_view(views, nv, slot);
    const ViewGPU& v = views[si.view];
    uint32_t local = slot - v.slot_base;
    uint32_t q = local / v.spp;
    si.sample = local - q * v.spp;
    uint32_t x, y;
    if (block_order(v)) {
        const uint32_t b = q >> 6, i = q & 63u, bw = v.tile_w >> 3;
        x = (b % bw) * 8u + (i & 7u);
        y = (b / bw) * 8u + (i >> 3);
    } else {
        y = q / v.tile_w;
        x = q - y * v.tile_w;
    }
    si.x = v.tile_x + x;
    si.y = v.tile_y + y;
    return si;
}

// generate_ray (path_trace_rgen.glsl:132-174); operation order fixed to match
// the CPU restatement bit for bit.
__device__ inline void generate_ray(const ViewGPU& v, uint32_t px, uint32_t py, RNG& rng, f3& o, f3& d) {
    float jx = next_float(rng);
    float jy = next_float(rng);
    float tcx = (((float)px + 0.5f) + jx) / (float)v.width;
    float tcy = (((float)py + 0.5f) + jy) / (float)v.height;
    float nx = tcx * 2.0f - 1.0f, ny = tcy * 2.0f - 1.0f;
    float tg[4];
#pragma unroll
    for (int r = 0; r < 4; r++) tg[r] = ((v.vpi[0 * 4 + r] * nx + v.vpi[1 * 4 + r] * ny) + v.vpi[2 * 4 + r] * 0.0f) + v.vpi[3 * 4 + r] * 1.0f;
    f3 target = mk(tg[0] / tg[3], tg[1] / tg[3], tg[2] / tg[3]);
    float angle = next_float(rng) * 2.0f * HH_PI;
    float radius = sqrtf(next_float(rng));
    float ox = cosf(angle) * radius * v.aperture, oy = sinf(angle) * radius * v.aperture;
    f3 cam = mk(v.cam[0], v.cam[1], v.cam[2]);
    f3 ap = (cam + mk(v.right[0], v.right[1], v.right[2]) * ox) + mk(v.up[0], v.up[1], v.up[2]) * oy;
    f3 rdir = -normalize(target - cam);
    f3 fp = mk(v.fplane[0], v.fplane[1], v.fplane[2]);
    float t = -(dot(cam, fp) + v.fplane[3]) / dot(rdir, fp);
    f3 focus = cam + rdir * t;
    o = ap;
    d = normalize(focus - ap);
}

__device__ inline float half_bits(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }

// Extension: equirectangular lookup, bilinear, repeating in x, clamped in y.
__device__ inline f3 sample_env_equirect(const SceneGPU& S, f3 d) {
    const int w = (int)S.env_eq_w, h = w / 2;
    const float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    const float phi = atan2f(d.x, -d.z);
    const float x = (phi / (2.0f * HH_PI) + 0.5f) * (float)w - 0.5f;
    const float y = (theta / HH_PI) * (float)h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float wx = x - fx, wy = y - fy;
    int i0 = ((int)fx % w + w) % w, i1 = (i0 + 1) % w;
    int j0 = min(max((int)fy, 0), h - 1), j1 = min(max((int)fy + 1, 0), h - 1);
    const uint16_t* t = S.env_eq;
    auto px = [&](int i, int j, int c) { return half_bits(t[4 * ((size_t)j * w + i) + c]); };
    return mk(lerpf(lerpf(px(i0, j0, 0), px(i1, j0, 0), wx), lerpf(px(i0, j1, 0), px(i1, j1, 0), wx), wy),
              lerpf(lerpf(px(i0, j0, 1), px(i1, j0, 1), wx), lerpf(px(i0, j1, 1), px(i1, j1, 1), wx), wy),
              lerpf(lerpf(px(i0, j0, 2), px(i1, j0, 2), wx), lerpf(px(i0, j1, 2), px(i1, j1, 2), wx), wy));
}

// Cube-map sampling: Vulkan face selection and seamless bilinear filtering
// (the reference's LINEAR env sampler, vk.cpp:3557-3587). A bilinear tap that
// falls off its face (i or j = -1 or n) reads the texel of the adjacent face
// that holds that tap's centre: the centre (2i+1-n, 2j+1-n)/n is extended on
// the face's plane and the face re-selected by its major axis, in exact
// integer arithmetic. A tap off a corner is the mean of the three texels
// that meet there. Texel centres lie half a texel from any texel border, so
// integer floors and the float computation agree with the CPU restatement.
__device__ inline int cube_select_i(int x, int y, int z, int& sc, int& tc, int& ma) {
    const int ax = abs(x), ay = abs(y), az = abs(z);
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (x >= 0) { sc = -z; tc = -y; return 0; }
        sc = z; tc = -y; return 1;
    }
    if (ay >= az) {
        ma = ay;
        if (y >= 0) { sc = x; tc = z; return 2; }
        sc = x; tc = -z; return 3;
    }
    ma = az;
    if (z >= 0) { sc = x; tc = -y; return 4; }
    sc = -x; tc = -y; return 5;
}
// (face, i, j) with i or j one texel outside [0, n): the adjacent face's
// texel index (face * n + j) * n + i.
__device__ inline size_t cube_wrap(int face, int i, int j, int n) {
    const int a = 2 * i + 1 - n, b = 2 * j + 1 - n;
    int x, y, z;
    switch (face) {  // inverse of the face selection: sc = a/n, tc = b/n
        case 0: x = n; y = -b; z = -a; break;
        case 1: x = -n; y = -b; z = a; break;
        case 2: x = a; y = n; z = b; break;
        case 3: x = a; y = -n; z = -b; break;
        case 4: x = a; y = -b; z = n; break;
        default: x = -a; y = -b; z = -n; break;
    }
    int sc, tc, ma;
    const int f2 = cube_select_i(x, y, z, sc, tc, ma);
    const int i2 = min(max((n * (sc + ma)) / (2 * ma), 0), n - 1);
    const int j2 = min(max((n * (tc + ma)) / (2 * ma), 0), n - 1);
    return ((size_t)f2 * n + j2) * n + i2;
}
__device__ inline float4 cube_texel(const float4* env, int n, int face, int i, int j) {
    const bool oi = (unsigned)i >= (unsigned)n, oj = (unsigned)j >= (unsigned)n;
    if (!oi && !oj) return env[((size_t)face * n + j) * n + i];
    const int ci = min(max(i, 0), n - 1), cj = min(max(j, 0), n - 1);
    if (oi && oj) {
        const float4 c0 = env[((size_t)face * n + cj) * n + ci];
        const float4 c1 = env[cube_wrap(face, i, cj, n)];
        const float4 c2 = env[cube_wrap(face, ci, j, n)];
        return make_float4(((c0.x + c1.x) + c2.x) / 3.0f, ((c0.y + c1.y) + c2.y) / 3.0f, ((c0.z + c1.z) + c2.z) / 3.0f,
                           0.0f);
    }
    return env[cube_wrap(face, i, j, n)];
}
__device__ inline f3 sample_env(const SceneGPU& S, f3 d) {
    if (S.env_eq_w > 0u) return sample_env_equirect(S, d);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d.x >= 0.0f) { face = 0; sc = -d.z; tc = -d.y; } else { face = 1; sc = d.z; tc = -d.y; }
    } else if (ay >= az) {
        ma = ay;
        if (d.y >= 0.0f) { face = 2; sc = d.x; tc = d.z; } else { face = 3; sc = d.x; tc = -d.z; }
    } else {
        ma = az;
        if (d.z >= 0.0f) { face = 4; sc = d.x; tc = -d.y; } else { face = 5; sc = -d.x; tc = -d.y; }
    }
    if (!(ma > 0.0f)) return mk(0.f, 0.f, 0.f);
    const int n = (int)S.env_size;
    float s = 0.5f * (sc / ma + 1.0f), t = 0.5f * (tc / ma + 1.0f);
    float x = s * (float)n - 0.5f, y = t * (float)n - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float wx = x - fx, wy = y - fy;
    const int i0 = (int)fx, j0 = (int)fy;
    const float4* env = S.env;
    float4 p00, p10, p01, p11;
    if (i0 >= 0 && j0 >= 0 && i0 + 1 < n && j0 + 1 < n) {   // interior: one face
        const float4* base = env + ((size_t)face * n + j0) * n + i0;
        p00 = base[0];
        p10 = base[1];
        p01 = base[n];
        p11 = base[n + 1];
    } else {
        p00 = cube_texel(env, n, face, i0, j0);
        p10 = cube_texel(env, n, face, i0 + 1, j0);
        p01 = cube_texel(env, n, face, i0, j0 + 1);
        p11 = cube_texel(env, n, face, i0 + 1, j0 + 1);
    }
    return mk(lerpf(lerpf(p00.x, p10.x, wx), lerpf(p01.x, p11.x, wx), wy),
              lerpf(lerpf(p00.y, p10.y, wx), lerpf(p01.y, p11.y, wx), wy),
              lerpf(lerpf(p00.z, p10.z, wx), lerpf(p01.z, p11.z, wx), wy));
}

__device__ inline f3 m4_point(const float* m, f3 v, float w) {
    return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * w,
              m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * w);
}
__device__ inline f3 m3_mul(const float* m, f3 v) {
    return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z, m[2] * v.x + m[6] * v.y + m[10] * v.z);
}

// populate_surface_properties (path_trace_rchit.glsl:256-280)
// Positions come from the traversal's triangle record and normals / UVs from
// the shading record of the same triangle (the vertex values themselves, so
// the interpolation is the vertex path's, bit for bit); only normal-mapped
// materials read the vertices, for their tangent frames.
__device__ inline void populate(const SceneGPU& S, uint32_t inst, uint32_t tri, float u, float v, Surface& p) {
    const float4* tw = S.tris + 3 * (size_t)tri;
    const float4* sr = S.shade + 4 * (size_t)tri;
    const float4 a = tw[0], b = tw[1], c = tw[2];
    const float4 r0 = sr[0], r1 = sr[1], r2 = sr[2], r3 = sr[3];
    uint32_t prim = __float_as_uint(a.w);
    uint32_t geom = __float_as_uint(b.w) & 0x7FFFFFFFu;
    const InstanceData& I = S.inst[inst];
    uint2 info = S.submesh_info[S.submesh_off[inst] + geom];
    const MaterialData& m = S.mats[info.y];
    float b0 = 1.0f - u - v, b1 = u, b2 = v;
    f3 pos = ld3(a) * b0 + ld3(b) * b1 + ld3(c) * b2;
    float uv0 = r0.x * b0 + r0.z * b1 + r1.x * b2;
    float uv1 = r0.y * b0 + r0.w * b1 + r1.y * b2;
    f3 n = normalize(mk(r1.z, r1.w, r2.x) * b0 + mk(r2.y, r2.z, r2.w) * b1 + ld3(r3) * b2);
    p.pos = m4_point(I.model, pos, 1.0f);
    p.vnormal = m3_mul(I.normal, n);
    float4 alb;
    if (m.tex0[0] == -1) alb = make_float4(m.albedo[0], m.albedo[1], m.albedo[2], m.albedo[3]);
    else alb = sample_tex(S, m.tex0[0], uv0, uv1);
    p.albedo = mk(alb.x, alb.y, alb.z);
    p.alpha_a = alb.w;
    if (m.tex0[1] == -1) {
        p.normal = p.vnormal;
    } else {
        const uint32_t mesh = S.inst_trav[inst].mesh;
        const uint32_t* ix = S.idx + S.ibase[mesh] + 3 * (prim + info.x);
        const float4* vb = S.vtx + 5 * (size_t)S.vbase[mesh];
        const float4* V0 = vb + 5 * (size_t)ix[0];
        const float4* V1 = vb + 5 * (size_t)ix[1];
        const float4* V2 = vb + 5 * (size_t)ix[2];
        f3 tg = normalize(ld3(V0[3]) * b0 + ld3(V1[3]) * b1 + ld3(V2[3]) * b2);
        f3 bt = normalize(ld3(V0[4]) * b0 + ld3(V1[4]) * b1 + ld3(V2[4]) * b2);
        f3 T = normalize(m3_mul(I.normal, tg)), B = normalize(m3_mul(I.normal, bt)), N = normalize(p.vnormal);
        float4 c = sample_tex(S, m.tex0[1], uv0, uv1);
        f3 tn = normalize(mk(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f));
        p.normal = normalize(T * tn.x + B * tn.y + N * tn.z);
    }
    auto chan = [](float4 c, int k) { k &= 3; return k == 0 ? c.x : (k == 1 ? c.y : (k == 2 ? c.z : c.w)); };
    p.roughness = m.tex0[2] == -1 ? m.roughness_metallic[0] : chan(sample_tex(S, m.tex0[2], uv0, uv1), m.tex1[2]);
    p.metallic = m.tex0[3] == -1 ? m.roughness_metallic[1] : chan(sample_tex(S, m.tex0[3], uv0, uv1), m.tex1[3]);
    if (m.tex1[0] == -1) p.emissive = mk(m.emissive[0], m.emissive[1], m.emissive[2]);
    else {
        float4 c = sample_tex(S, m.tex1[0], uv0, uv1);
        p.emissive = mk(c.x, c.y, c.z);
    }
    p.roughness = fmaxf(p.roughness, HH_MIN_ROUGHNESS);
    p.F0 = mk(0.03f, 0.03f, 0.03f) * (1.0f - p.metallic) + p.albedo * p.metallic;
}

// One NEE light sample (path_trace_rchit.glsl:284-451, 455-483). Returns the
// contribution that becomes radiance if the shadow ray is unoccluded, and the
// shadow ray. want_ray = false when the contribution is exactly zero.
struct ShadowRay {
    f3 o, d;
    float tmin, tmax;
    f3 contrib;
    uint32_t flags;
    bool want;
};
__device__ inline ShadowRay direct_lighting(const SceneGPU& S, const Surface& p, f3 ray_dir, f3 T, uint32_t depth,
                                            float shadow_bias, uint32_t vflags, RNG& rng) {
    ShadowRay sr;
    sr.want = false;
    sr.flags = depth == 0 ? 0u : (HH_TRACE_OPAQUE_BIT | HH_TRACE_TERMINATE_BIT);
    sr.tmin = 0.0001f;
    sr.tmax = 10000.0f;
    sr.o = p.pos + p.normal * shadow_bias;
    // Reference: uniform pick among the packed lights. Power mode (extension):
    // uniform among [all emitters as one entry] + the lights after the
    // area-light records; type 5 = an emitter chosen by power.
    const bool power = (vflags & HH_FLAG_LIGHT_POWER_BIT) != 0u;
    const uint32_t nemit = power && S.num_emit > 0u ? 1u : 0u;
    const uint32_t nl = power ? S.num_lights - S.num_area + nemit : S.num_lights;
    uint32_t li = next_uint(rng, nl);
    if (nl == 0) return sr;
    const bool emitter = nemit != 0u && li == 0u;
    const LightData& L = S.lights[emitter ? 0u : (power ? li - nemit + S.num_area : li)];
    uint32_t type = emitter ? 5u : (uint32_t)L.d0[0];
    f3 Wi = mk(0.f, 0.f, 0.f), Li = mk(0.f, 0.f, 0.f);
    float pdf = 0.0f;
    if (type == 0u) {
        float r0 = next_float(rng);
        float r1 = next_float(rng);
        f3 ld = -mk(L.d1[0], L.d1[1], L.d1[2]);
        f3 lt = normalize(cross(ld, mk(0.f, 1.f, 0.f)));
        f3 lb = normalize(cross(lt, ld));
        float pr = L.d2[3] * sqrtf(r0), pa = r1 * 2.0f * HH_PI;
        Wi = normalize(ld + lt * (pr * cosf(pa)) + lb * (pr * sinf(pa)));
        Li = mk(L.d0[1], L.d0[2], L.d0[3]) * L.d1[3];
    } else if (type == 1u || type == 2u) {
        float r0 = next_float(rng);
        float r1 = next_float(rng);
        f3 to = mk(L.d2[0], L.d2[1], L.d2[2]) - p.pos;
        f3 ld = normalize(to);
        float dist = length(to);
        float lr = L.d2[3] / dist;
        f3 lt = normalize(cross(ld, mk(0.f, 1.f, 0.f)));
        f3 lb = normalize(cross(lt, ld));
        float pr = lr * sqrtf(r0), pa = r1 * 2.0f * HH_PI;
        Wi = normalize(ld + lt * (pr * cosf(pa)) + lb * (pr * sinf(pa)));
        if (type == 1u) {
            float x = dot(ld, -mk(L.d1[0], L.d1[1], L.d1[2]));
            float tt = clampf((x - L.d3[1]) / (L.d3[0] - L.d3[1]), 0.0f, 1.0f);
            float att = tt * tt * (3.0f - 2.0f * tt);
            Li = mk(L.d0[1], L.d0[2], L.d0[3]) * L.d1[3] * att / (dist * dist);
        } else {
            Li = mk(L.d0[1], L.d0[2], L.d0[3]) * L.d1[3] / (dist * dist);
        }
        sr.tmax = dist;
    } else if (type == 3u) {
        float r0 = next_float(rng);
        float r1 = next_float(rng);
        Wi = sample_cosine_lobe(p.normal, r0, r1);
        Li = sample_env(S, Wi);
        pdf = dot(p.normal, Wi) / HH_PI;
    } else if (type == 4u) {
        uint32_t mesh_node = (uint32_t)L.d0[1];
        uint32_t prim = next_uint(rng, (uint32_t)L.d1[2]);  // reference reads .z (= 0)
        uint32_t mat_idx = (uint32_t)L.d0[2], off = (uint32_t)L.d0[3];
        const InstanceData& I = S.inst[mesh_node];
        uint32_t mesh = S.inst_trav[mesh_node].mesh;
        uint32_t pid = prim + off;
        const uint32_t* ix = S.idx + S.ibase[mesh] + 3 * pid;
        const float4* vb = S.vtx + 5 * (size_t)S.vbase[mesh];
        float b0 = next_float(rng);
        float b1 = next_float(rng);
        float su0 = sqrtf(b0);
        float bx = 1.0f - su0, by = b1 * su0;
        float4 q0 = vb[5 * (size_t)ix[0]], q1 = vb[5 * (size_t)ix[1]], q2 = vb[5 * (size_t)ix[2]];
        f3 P0 = m4_point(I.model, ld3(q0), q0.w), P1 = m4_point(I.model, ld3(q1), q1.w), P2 = m4_point(I.model, ld3(q2), q2.w);
        float w0 = 1.0f - bx - by;
        f3 lp = P0 * w0 + P1 * bx + P2 * by;
        f3 ln = normalize(m3_mul(I.normal, ld3(vb[5 * (size_t)ix[0] + 2]) * w0 + ld3(vb[5 * (size_t)ix[1] + 2]) * bx +
                                                ld3(vb[5 * (size_t)ix[2] + 2]) * by));
        f3 ldir = p.pos - lp;
        float dsq = dot(ldir, ldir);
        float area = 0.5f * length(cross(P1 - P0, P2 - P0));
        if (area == 0.0f || dsq == 0.0f) return sr;
        float dist = sqrtf(dsq);
        ldir = ldir / dist;
        sr.tmax = fmaxf(0.0f, dist - HH_EPS);
        float ct = dot(ln, ldir);
        if (ct == 0.0f) return sr;
        const MaterialData& lm = S.mats[mat_idx];
        Li = mk(lm.emissive[0], lm.emissive[1], lm.emissive[2]);
        Wi = -ldir;
        pdf = dsq / fmaxf(HH_EPS, ct * area);
    } else if (type == 5u) {
        // emitter triangle by power (inverse CDF), uniform point on it;
        // pdf in solid angle = P(tri) / area * d^2 / cos (front face only)
        const float u = next_float(rng);
        uint32_t lo = 0, hi = S.num_emit - 1u;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (u < S.emit_cdf[mid]) hi = mid;
            else lo = mid + 1u;
        }
        const float4 a = S.emit_tri[3 * (size_t)lo], b = S.emit_tri[3 * (size_t)lo + 1], c3 = S.emit_tri[3 * (size_t)lo + 2];
        const float4 le = S.emit_le[lo];
        float b0 = next_float(rng);
        float b1 = next_float(rng);
        float su0 = sqrtf(b0);
        float bx = 1.0f - su0, by = b1 * su0;
        f3 P0 = ld3(a), P1 = ld3(b), P2 = ld3(c3);
        float w0 = 1.0f - bx - by;
        f3 lp = P0 * w0 + P1 * bx + P2 * by;
        f3 ln = normalize(cross(P1 - P0, P2 - P0));
        f3 ldir = p.pos - lp;
        float dsq = dot(ldir, ldir);
        if (dsq == 0.0f) return sr;
        float dist = sqrtf(dsq);
        ldir = ldir / dist;
        sr.tmax = fmaxf(0.0f, dist - HH_EPS);
        float ct = dot(ln, ldir);
        if (!(ct > 0.0f)) return sr;
        Li = mk(le.x, le.y, le.z);
        Wi = -ldir;
        pdf = le.w / a.w * dsq / ct;
    }
    f3 Wo = -ray_dir;
    f3 Wh = normalize(Wo + Wi);
    f3 brdf = evaluate_uber(p, Wo, Wh, Wi);
    float ct = clampf(dot(p.normal, Wi), 0.0f, 1.0f);
    if (is_black(Li)) return sr;
    f3 c = T * brdf * ct * Li;
    if (pdf != 0.0f) c = c / pdf;
    c = c * (float)nl;
    sr.contrib = c;
    sr.d = Wi;
    sr.want = !is_black(c);
    return sr;
}

// First-hit features of one sample (denoiser inputs): albedo (the clamped
// environment radiance on a miss), depth as 1/(1+t) (0 on a miss), shading
// normal (0 on a miss).
__device__ inline void write_features(float4* feat, uint32_t stride, uint32_t slot, f3 albedo, float t, f3 n, bool miss) {
    if (miss) albedo = mk(fminf(albedo.x, 1.0f), fminf(albedo.y, 1.0f), fminf(albedo.z, 1.0f));
    feat[slot] = make_float4(albedo.x, albedo.y, albedo.z, miss ? 0.0f : 1.0f / (1.0f + t));
    feat[stride + slot] = make_float4(n.x, n.y, n.z, 0.0f);
}

__device__ inline float luminance(float3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
template <int BS>
__global__ void __launch_bounds__(BS) k_generate(const ViewGPU* __restrict__ views, uint32_t nv, uint32_t total,
                                                 float4* __restrict__ q_o, float4* __restrict__ q_d,
                                                 uint32_t* __restrict__ q_slot, float4* __restrict__ L,
                                                 uint2* __restrict__ rngs, uint32_t* __restrict__ counters, uint32_t seg,
                                                 uint32_t clear_radiance) {
    // The first extension queue is dealt out to the segments in 64-entry
    // chunks: chunk c goes to segment c % kQShards at offset (c / kQShards) * 64.
    const uint32_t stride = gridDim.x * BS;
    for (uint32_t slot = blockIdx.x * BS + threadIdx.x; slot < total; slot += stride) {
        SlotInfo si = slot_info(views, nv, slot);
        const ViewGPU& v = views[si.view];
        RNG rng = rng_init(si.x, si.y, v.first_frame + si.sample);
        f3 o, d;
        generate_ray(v, si.x, si.y, rng, o, d);
        const uint32_t c = slot >> 6;
        const uint32_t qi = (c % kQShards) * seg + (c / kQShards) * 64u + (slot & 63u);
        q_o[qi] = make_float4(o.x, o.y, o.z, 0.001f);
        q_d[qi] = make_float4(d.x, d.y, d.z, 10000.0f);
        q_slot[qi] = slot;
        // the first shade launch starts every path from T = 1, L = 0 without
        // reading them; only a frame without bounces leaves L to this kernel
        if (clear_radiance) L[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        rngs[slot] = make_uint2(rng.s0, rng.s1);
    }
    if (blockIdx.x == 0 && threadIdx.x < kQShards) {
        const uint32_t sh = threadIdx.x, nch = (total + 63u) >> 6;
        uint32_t items = (nch / kQShards + (sh < nch % kQShards ? 1u : 0u)) * 64u;
        if (nch > 0u && (nch - 1u) % kQShards == sh) items -= nch * 64u - total;
        counters[ctr_index(kCtrExt0, sh)] = items;
        counters[ctr_index(kCtrExt1, sh)] = 0u;
        counters[ctr_index(kCtrShadow, sh)] = 0u;
        counters[ctr_index(kCtrExtHead, sh)] = 0u;
        counters[ctr_index(kCtrShadowHead, sh)] = 0u;
    }
}

// Adds per-lane traversal counts into three device counters (one atomic per wave).
__device__ inline void flush_counts(const TravCount& tc, unsigned long long* out) {
    unsigned long long a = tc.nodes, b = tc.tris, c = tc.insts;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
    if (lane_id() == 0) {
        atomicAdd(&out[0], a);
        atomicAdd(&out[1], b);
        atomicAdd(&out[2], c);
    }
}

// Persistent queue traversal: every wave keeps up to 64 rays in flight and,
// whenever at least `refill` of its lanes have finished (or all have), claims
// the next rays for them with one atomic on the read head of a queue segment.
// A wave starts on its own segment and moves on to the following ones when a
// segment runs dry, so the claims of the grid spread over kQShards counters.
// Lanes step their rays independently, so a wave is not held by its longest
// ray. (Claiming larger private chunks per atomic was measured slower: it
// starves waves of work when the deep-bounce queues are short.)
struct QueueCursor {
    const uint32_t* count;  // segment lengths, kCtrStride apart
    uint32_t* head;         // segment read heads, kCtrStride apart
    uint32_t seg, refill;
    uint32_t shard, tried;  // wave-uniform
    bool drained;           // every segment is exhausted
    __device__ inline QueueCursor(const uint32_t* count_, uint32_t* head_, uint32_t seg_, uint32_t refill_)
        : count(count_), head(head_), seg(seg_), refill(refill_), shard(wave_shard()), tried(0u), drained(false) {}
    // Queue entry for this lane (kNoEntry: none). Called by the whole wave in
    // converged control flow.
    __device__ inline uint32_t next(bool idle_lane) {
        if (drained) return kNoEntry;
        const unsigned long long idle = __ballot(idle_lane);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle < refill && nidle != 64u) return kNoEntry;
        // idle lanes below this one (mbcnt: no lane mask or lane index kept
        // in registers); the lowest idle lane (rank 0) claims for the wave
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
        uint32_t got = kNoEntry, done = 0u;
        while (done < nidle) {
            const uint32_t want = nidle - done;
            const uint32_t n = count[shard * kCtrStride];
            uint32_t base = 0;
            if (idle_lane && rank == 0u) {
                // a segment already read to its end costs no atomic
                base = __hip_atomic_load(head + shard * kCtrStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (base < n) base = atomicAdd(head + shard * kCtrStride, want);
            }
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, (int)__builtin_ctzll(idle));
            const uint32_t avail = base < n ? min(n - base, want) : 0u;
            if (idle_lane && rank >= done && rank < done + avail) got = shard * seg + base + (rank - done);
            done += avail;
            if (base + want >= n) {
                shard = shard + 1u == kQShards ? 0u : shard + 1u;
                if (++tried == kQShards) {
                    drained = true;
                    break;
                }
            }
        }
        return got;
    }
};

// Sum of the kQShards segment lengths.
__device__ inline uint32_t queue_total(const uint32_t* count) {
    uint32_t n = 0;
    for (uint32_t k = 0; k < kQShards; k++) n += count[k * kCtrStride];
    return n;
}

// Five waves per SIMD (at most 102 VGPRs, no spills): measured 0.470 -> 0.452
// ms of extension per 1080p frame against six waves with a few spills
// (k_shadow, whose lanes finish at their first hit, is faster at six).
template <int LY, int BS, int MODE, bool COUNT>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(5))) k_extend(const SceneGPU* __restrict__ scenes, const ViewGPU* __restrict__ views, uint32_t nv,
                                               const float4* __restrict__ q_o, const float4* __restrict__ q_d,
                                               const uint32_t* __restrict__ q_slot, const uint32_t* __restrict__ count,
                                               uint32_t* __restrict__ head, uint32_t seg, uint32_t refill, uint32_t rflags,
                                               float4* __restrict__ hit, uint32_t* __restrict__ hit_tri,
                                               unsigned long long* __restrict__ stats, unsigned long long* __restrict__ tcounts,
                                               uint32_t* __restrict__ spill) {
    __shared__ uint32_t stack[kLdsStack * BS];
    const LaneStack stk{to_lds(&stack[threadIdx.x]), BS, to_global(spill + blockIdx.x * BS + threadIdx.x), gridDim.x * BS};
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[0], (unsigned long long)queue_total(count));
    QueueCursor qc(count, head, seg, refill);
    TravCount tc{0u, 0u, 0u};
    Tracer<LY, false, COUNT> tr;
    uint32_t q = 0;
    bool active = false;
    for (;;) {
        const uint32_t nq = qc.next(!active);
        if (nq != kNoEntry) {
            const uint32_t slot = q_slot[nq];
            const uint32_t view = nv == 1 ? 0u : find_view(views, nv, slot);
            const float4 o = q_o[nq], d = q_d[nq];
            tr.begin(scenes[views[view].scene], ld3(o), ld3(d), o.w, d.w, rflags, &tc);
            q = nq;
            active = true;
        }
        if (__ballot(active) == 0ull) break;
        if (active && tr.template step<MODE>(stk)) {
            const HitRec& h = tr.h;
            hit[q] = make_float4(h.found ? h.t : -1.0f, h.u, h.v, __uint_as_float(h.inst));
            hit_tri[q] = h.found ? h.tri : 0xFFFFFFFFu;
            active = false;
        }
    }
    if (COUNT) flush_counts(tc, tcounts);
}

// Packets pay off once a wave holds samples of few pixels: measured per 1080p
// frame of extension, 0.97 vs 0.99 ms at 4 samples per pixel in flight, 0.58 vs
// 0.67 at 16; at one sample per pixel (64 pixels of a row) the single-frame
// path was slower with them (3.46 vs 3.16 ms).
constexpr uint32_t kPacketMinSpp = 4;

// Camera rays in packets (trace_packet): a wave claims 64 consecutive entries
// of the first extension queue, i.e. jittered samples of one pixel or
// neighbouring pixels (pixel-major slots, 64-entry chunks), and traces them
// together. Used when every view of the frame renders the same scene.
template <int LY, int BS>
__global__ void __launch_bounds__(BS) k_extend_packet(const SceneGPU* __restrict__ scenes, const ViewGPU* __restrict__ views,
                                                      const float4* __restrict__ q_o, const float4* __restrict__ q_d,
                                                      const uint32_t* __restrict__ count, uint32_t* __restrict__ head,
                                                      uint32_t seg, uint32_t rflags, float4* __restrict__ hit,
                                                      uint32_t* __restrict__ hit_tri, unsigned long long* __restrict__ stats) {
    __shared__ uint32_t pstack[kPacketStack * (BS / 64)];
    HH_LDS uint32_t* const stack = to_lds(&pstack[(threadIdx.x >> 6) * kPacketStack]);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[0], (unsigned long long)queue_total(count));
    const SceneGPU& S = scenes[views[0].scene];
    QueueCursor qc(count, head, seg, 64u);
    for (;;) {
        const uint32_t q = qc.next(true);
        const bool active = q != kNoEntry;
        if (__ballot(active) == 0ull) break;
        f3 o = mk(0.f, 0.f, 0.f), d = mk(0.f, 0.f, 1.f);
        float t0 = 0.0f, t1 = 0.0f;
        if (active) {
            const float4 a = q_o[q], b = q_d[q];
            o = ld3(a);
            t0 = a.w;
            d = ld3(b);
            t1 = b.w;
        }
        HitRec h;
        trace_packet<LY>(S, active, o, d, t0, t1, rflags, h, stack);
        if (active) {
            hit[q] = make_float4(h.found ? h.t : -1.0f, h.u, h.v, __uint_as_float(h.inst));
            hit_tri[q] = h.found ? h.tri : 0xFFFFFFFFu;
        }
    }
}

// Four waves per SIMD (at most 128 VGPRs, no spills): 0.229 -> 0.196 ms per
// 1080p frame over the unconstrained three; five or six waves spill and are
// slower.
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) k_shade(const SceneGPU* __restrict__ scenes, const ViewGPU* __restrict__ views, uint32_t nv,
                                              uint32_t depth, const float4* __restrict__ q_o, const float4* __restrict__ q_d,
                                              const uint32_t* __restrict__ q_slot, const uint32_t* __restrict__ count,
                                              uint32_t seg, const float4* __restrict__ hit, const uint32_t* __restrict__ hit_tri,
                                              float4* __restrict__ T, float4* __restrict__ L, uint2* __restrict__ rngs,
                                              float4* __restrict__ n_o, float4* __restrict__ n_d, uint32_t* __restrict__ n_slot,
                                              uint32_t* __restrict__ n_count, float4* __restrict__ s_o, float4* __restrict__ s_d,
                                              float4* __restrict__ s_c, uint32_t* __restrict__ s_slot,
                                              uint32_t* __restrict__ s_count, float4* __restrict__ feat, uint32_t feat_stride) {
    // The input queue is read as one list over its segments: entry q lives in
    // segment k with first[k] <= q < first[k] + count[k] (first = exclusive
    // prefix sum of the lengths, held in lanes 0..kQShards-1).
    const uint32_t lane = __lane_id();
    const uint32_t len = lane < kQShards ? count[lane * kCtrStride] : 0u;
    uint32_t incl = len;
    for (uint32_t d = 1; d < kQShards; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    const uint32_t first = incl - len;
    const uint32_t n = __shfl(incl, (int)kQShards - 1);
    const uint32_t out_shard = wave_shard();
    uint32_t* const n_ctr = n_count + out_shard * kCtrStride;
    uint32_t* const s_ctr = s_count + out_shard * kCtrStride;
    const uint32_t stride = gridDim.x * BS;
    // Keep whole waves in the loop so the ballot-based compaction sees every
    // active lane: iterate over the count rounded up to the grid stride.
    const uint32_t n_round = ((n + stride - 1) / stride) * stride;
    for (uint32_t ql = blockIdx.x * BS + threadIdx.x; ql < n_round; ql += stride) {
        const bool active = ql < n;
        uint32_t q = ql, qk = 0;
#pragma unroll
        for (uint32_t k = 1; k < kQShards; k++) {
            const uint32_t f = (uint32_t)__builtin_amdgcn_readlane((int)first, (int)k);
            if (ql >= f) {
                qk = k;
                q = ql - f;
            }
        }
        q += qk * seg;
        bool want_ext = false, want_sh = false;
        f3 ext_o = mk(0.f, 0.f, 0.f), ext_d = mk(0.f, 0.f, 1.f);
        ShadowRay sr;
        sr.want = false;
        uint32_t slot = 0;
        if (active) {
            slot = q_slot[q];
            uint32_t view = nv == 1 ? 0u : find_view(views, nv, slot);
            const ViewGPU& V = views[view];
            const SceneGPU& S = scenes[V.scene];
            float4 hh = hit[q];
            f3 dir = ld3(q_d[q]);
            float4 T4 = depth == 0 ? make_float4(1.f, 1.f, 1.f, 0.f) : T[slot];
            f3 Tp = ld3(T4);
            float4 L4 = depth == 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : L[slot];
            f3 Lp = ld3(L4);
            uint2 rs = rngs[slot];
            RNG rng{rs.x, rs.y};
            const bool features = depth == 0 && V.aux != nullptr;
            if (hh.x < 0.0f) {
                // miss (path_trace_rmiss.glsl:60-65)
                f3 e = sample_env(S, dir);
                Lp = Lp + (depth == 0 ? e : Tp * e);
                if (features) write_features(feat, feat_stride, slot, e, 0.0f, mk(0.f, 0.f, 0.f), true);
            } else {
                Surface p;
                populate(S, __float_as_uint(hh.w), hit_tri[q], hh.y, hh.z, p);
                if (features) write_features(feat, feat_stride, slot, p.albedo, hh.x, p.normal, false);
                if (depth == 0 && !is_black(p.emissive)) Lp = Lp + p.emissive;
                sr = direct_lighting(S, p, dir, Tp, depth, V.shadow_bias, V.flags, rng);
                if ((V.flags & HH_FLAG_DIRECT_ONLY_BIT) == 0u && depth + 1 < V.max_bounces) {
                    f3 Wo = -dir, Wi;
                    float pdf;
                    f3 brdf = sample_uber(p, Wo, rng, Wi, pdf);
                    float ct = clampf(dot(p.normal, Wi), 0.0f, 1.0f);
                    f3 Tn = (Tp * (brdf * ct)) / pdf;
                    float prob = fmaxf(Tn.x, fmaxf(Tn.y, Tn.z));
                    // the RR draw is consumed at every bounce (stream
                    // alignment); before rr_start it neither ends nor rescales
                    float rr = next_float(rng);
                    const bool roulette = depth >= V.rr_start;
                    if (!roulette || !(rr > prob)) {
                        if (roulette) Tn = Tn * (1.0f / prob);
                        want_ext = true;
                        ext_o = p.pos;
                        ext_d = Wi;
                        T[slot] = make_float4(Tn.x, Tn.y, Tn.z, 0.f);
                    }
                }
                rngs[slot] = make_uint2(rng.s0, rng.s1);
            }
            L[slot] = make_float4(Lp.x, Lp.y, Lp.z, 0.f);
            want_sh = sr.want;
        }
        const uint32_t ei = out_shard * seg + wave_alloc_keyed<8>(n_ctr, want_ext, octant(ext_d));
        if (want_ext) {
            n_o[ei] = make_float4(ext_o.x, ext_o.y, ext_o.z, 0.0001f);
            n_d[ei] = make_float4(ext_d.x, ext_d.y, ext_d.z, 10000.0f);
            n_slot[ei] = slot;
        }
        const uint32_t si = out_shard * seg + wave_alloc_keyed<8>(s_ctr, want_sh, octant(sr.d));
        if (want_sh) {
            s_o[si] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tmin);
            s_d[si] = make_float4(sr.d.x, sr.d.y, sr.d.z, sr.tmax);
            s_c[si] = make_float4(sr.contrib.x, sr.contrib.y, sr.contrib.z, __uint_as_float(sr.flags));
            s_slot[si] = slot;
        }
    }
}

template <int LY, int BS, int MODE, bool COUNT>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(6))) k_shadow(const SceneGPU* __restrict__ scenes, const ViewGPU* __restrict__ views, uint32_t nv,
                                               const float4* __restrict__ s_o, const float4* __restrict__ s_d,
                                               const float4* __restrict__ s_c, const uint32_t* __restrict__ s_slot,
                                               const uint32_t* __restrict__ count, uint32_t* __restrict__ head,
                                               uint32_t seg, uint32_t refill, float4* __restrict__ L,
                                               unsigned long long* __restrict__ stats, unsigned long long* __restrict__ tcounts,
                                               uint32_t* __restrict__ spill) {
    __shared__ uint32_t stack[kLdsStack * BS];
    const LaneStack stk{to_lds(&stack[threadIdx.x]), BS, to_global(spill + blockIdx.x * BS + threadIdx.x), gridDim.x * BS};
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[1], (unsigned long long)queue_total(count));
    QueueCursor qc(count, head, seg, refill);
    TravCount tc{0u, 0u, 0u};
    Tracer<LY, true, COUNT> tr;
    uint32_t q = 0;
    bool active = false;
    for (;;) {
        const uint32_t nq = qc.next(!active);
        if (nq != kNoEntry) {
            const uint32_t slot = s_slot[nq];
            const uint32_t view = nv == 1 ? 0u : find_view(views, nv, slot);
            const float4 o = s_o[nq], d = s_d[nq];
            const uint32_t rflags = __float_as_uint(s_c[nq].w);
            tr.begin(scenes[views[view].scene], ld3(o), ld3(d), o.w, d.w, rflags, &tc);
            q = nq;
            active = true;
        }
        if (__ballot(active) == 0ull) break;
        if (active && tr.template step<MODE>(stk)) {
            if (!tr.h.found) {
                // one shadow ray per path slot per bounce: no other lane adds to L[slot]
                const uint32_t slot = s_slot[q];
                const float4 c = s_c[q], l = L[slot];
                L[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.f);
            }
            active = false;
        }
    }
    if (COUNT) flush_counts(tc, tcounts + 3);
}

// Clears the next extension queue, the shadow queue and both read heads.
__global__ void k_reset_queues(uint32_t* counters, uint32_t next_ext) {
    const uint32_t sh = threadIdx.x;
    if (sh < kQShards) {
        counters[ctr_index(next_ext, sh)] = 0u;
        counters[ctr_index(kCtrShadow, sh)] = 0u;
        counters[ctr_index(kCtrExtHead, sh)] = 0u;
        counters[ctr_index(kCtrShadowHead, sh)] = 0u;
    }
}

// path_trace_rgen.glsl:219-247: clamp and progressive mean, samples in order.
// The feature planes follow the same recurrence, so they average the same
// samples as the radiance (samples 1..f after frame f); M2 is Welford's sum
// of squared luminance deviations about that running mean.
// Samples of a pixel are consecutive slots (pixel-major), so a block stages
// kAccChunk samples of its kAccBlock pixels in LDS with contiguous loads, then
// each thread applies its pixel's samples in order from LDS.
constexpr uint32_t kAccBlock = 256;
constexpr uint32_t kAccChunk = 8;
__global__ void __launch_bounds__(kAccBlock) k_accumulate(const ViewGPU* __restrict__ views, uint32_t nv,
                                                          const float4* __restrict__ L, const float4* __restrict__ feat,
                                                          uint32_t feat_stride) {
    __shared__ float4 stage[kAccBlock * (kAccChunk + 1)];  // one padding entry per pixel row
    const ViewGPU& v = views[blockIdx.y];
    const uint32_t npix = v.tile_w * v.tile_h, spp = v.spp;
    float4* const aux = v.aux;
    // every loop bound below is uniform over the block, so all threads reach
    // every barrier
    for (uint32_t p0 = blockIdx.x * kAccBlock; p0 < npix; p0 += gridDim.x * kAccBlock) {
        const uint32_t p = p0 + threadIdx.x;
        const bool valid = p < npix;
        const uint32_t npb = min(kAccBlock, npix - p0);
        size_t pi = 0;
        float4 prev = make_float4(0.f, 0.f, 0.f, 0.f), a0 = prev, a1 = prev;
        if (valid) {
            const uint32_t y = v.tile_y + p / v.tile_w, x = v.tile_x + p % v.tile_w;
            pi = (size_t)y * v.width + x;
            prev = v.accum[pi];
            if (aux) {
                a0 = aux[pi];
                a1 = aux[(size_t)v.width * v.height + pi];
            }
        }
        for (uint32_t s0 = 0; s0 < spp; s0 += kAccChunk) {
            const uint32_t cn = min(kAccChunk, spp - s0);
            __syncthreads();  // the previous round's LDS reads are done
            for (uint32_t e = threadIdx.x; e < npb * cn; e += kAccBlock) {
                const uint32_t pp = e / cn, ss = e - pp * cn;
                stage[pp * (kAccChunk + 1) + ss] = L[v.slot_base + pixel_order(v, p0 + pp) * spp + s0 + ss];
            }
            __syncthreads();
            if (!valid) continue;
            for (uint32_t k = 0; k < cn; k++) {
                const uint32_t s = s0 + k;
                const uint32_t slot = v.slot_base + pixel_order(v, p) * spp + s;
                const float4 l = stage[threadIdx.x * (kAccChunk + 1) + k];
                float3 c;
                if (v.flags & HH_FLAG_NO_CLAMP_BIT) c = make_float3(l.x, l.y, l.z);
                else c = make_float3(fminf(l.x, 1.0f), fminf(l.y, 1.0f), fminf(l.z, 1.0f));
                uint32_t f = v.first_frame + s;
                const float lum_prev = luminance(make_float3(prev.x, prev.y, prev.z));
                if (f == 0) prev = make_float4(c.x, c.y, c.z, 1.0f);
                else {
                    float ff = (float)f;
                    prev = make_float4(prev.x + (c.x - prev.x) / ff, prev.y + (c.y - prev.y) / ff, prev.z + (c.z - prev.z) / ff, 1.0f);
                }
                if (aux) {
                    const float4 g0 = feat[slot], g1 = feat[feat_stride + slot];
                    const float xl = luminance(c);
                    if (f == 0) {
                        a0 = g0;
                        a1 = make_float4(g1.x, g1.y, g1.z, 0.0f);
                    } else {
                        const float ff = (float)f;
                        a0 = make_float4(a0.x + (g0.x - a0.x) / ff, a0.y + (g0.y - a0.y) / ff, a0.z + (g0.z - a0.z) / ff,
                                         a0.w + (g0.w - a0.w) / ff);
                        const float m2 = f == 1 ? 0.0f : a1.w + (xl - lum_prev) * (xl - luminance(make_float3(prev.x, prev.y, prev.z)));
                        a1 = make_float4(a1.x + (g1.x - a1.x) / ff, a1.y + (g1.y - a1.y) / ff, a1.z + (g1.z - a1.z) / ff, m2);
                    }
                }
            }
        }
        if (valid) {
            v.accum[pi] = prev;
            if (aux) {
                aux[pi] = a0;
                aux[(size_t)v.width * v.height + pi] = a1;
            }
        }
    }
}

// The bounces after the first of every path, in one launch: each lane takes a
// path from the depth-1 extension queue (ray, slot; throughput, radiance and
// RNG from the path state) and traces it to its end, then takes the next
// entry (claimed per wave like the traversal kernels do), so a wave is not
// held by its longest path. One loop iteration is one bounce (extension ray,
// shading, shadow ray). Four waves per SIMD (at most 128 VGPRs): one 1080p
// frame 3.17 -> 3.03 ms over the unconstrained three.
template <int W, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) k_megakernel(const SceneGPU* __restrict__ scenes, const ViewGPU* __restrict__ views,
                                                   uint32_t nv, const float4* __restrict__ q_o, const float4* __restrict__ q_d,
                                                   const uint32_t* __restrict__ q_slot, const uint32_t* __restrict__ count,
                                                   uint32_t* __restrict__ head, uint32_t seg, uint32_t refill,
                                                   const float4* __restrict__ Tin, const uint2* __restrict__ rngs,
                                                   float4* __restrict__ Lout, unsigned long long* __restrict__ stats,
                                                   uint32_t* __restrict__ spill, uint32_t depth0) {
    __shared__ uint32_t stack[kLdsStack * BS];
    const LaneStack stk{to_lds(&stack[threadIdx.x]), BS, to_global(spill + blockIdx.x * BS + threadIdx.x), gridDim.x * BS};
    QueueCursor qc(count, head, seg, refill);
    unsigned long long n_ext = 0, n_sh = 0;
    bool active = false;
    uint32_t slot = 0, view = 0, depth = 0, rflags = 0;
    f3 o = mk(0.f, 0.f, 0.f), d = mk(0.f, 0.f, 1.f), T = mk(1.f, 1.f, 1.f), Lp = mk(0.f, 0.f, 0.f);
    float tmin = 0.001f;
    RNG rng{0u, 0u};
    for (;;) {
        const uint32_t nq = qc.next(!active);
        if (nq != kNoEntry) {
            slot = q_slot[nq];
            view = nv == 1 ? 0u : find_view(views, nv, slot);
            const float4 o4 = q_o[nq];
            o = ld3(o4);
            tmin = o4.w;
            d = ld3(q_d[nq]);
            T = ld3(Tin[slot]);
            Lp = ld3(Lout[slot]);
            const uint2 rs = rngs[slot];
            rng = RNG{rs.x, rs.y};
            rflags = HH_TRACE_OPAQUE_BIT;
            depth = depth0;
            active = true;
        }
        if (__ballot(active) == 0ull) break;
        if (active) {
            const ViewGPU& V = views[view];
            const SceneGPU& S = scenes[V.scene];
            bool finish = false;
            HitRec h;
            n_ext++;
            trace_ray<W, false>(S, o, d, tmin, 10000.0f, rflags, h, stk);
            if (!h.found) {
                Lp = Lp + T * sample_env(S, d);
                finish = true;
            } else {
                Surface p;
                populate(S, h.inst, h.tri, h.u, h.v, p);
                const ShadowRay sr = direct_lighting(S, p, d, T, depth, V.shadow_bias, V.flags, rng);
                // The bounce is sampled before the shadow ray is traced (the
                // RNG draws keep their order: light sample, BRDF sample, RR),
                // so the surface is dead across that traversal.
                if ((V.flags & HH_FLAG_DIRECT_ONLY_BIT) != 0u || depth + 1 >= V.max_bounces) {
                    finish = true;
                } else {
                    f3 Wo = -d, Wi;
                    float pdf;
                    f3 brdf = sample_uber(p, Wo, rng, Wi, pdf);
                    float ct = clampf(dot(p.normal, Wi), 0.0f, 1.0f);
                    f3 Tn = (T * (brdf * ct)) / pdf;
                    float prob = fmaxf(Tn.x, fmaxf(Tn.y, Tn.z));
                    const float rr = next_float(rng);
                    if (depth >= V.rr_start && rr > prob) {
                        finish = true;
                    } else {
                        T = depth >= V.rr_start ? Tn * (1.0f / prob) : Tn;
                        o = p.pos;
                        d = Wi;
                        tmin = 0.0001f;
                        rflags = HH_TRACE_OPAQUE_BIT;
                        depth++;
                    }
                }
                if (sr.want) {
                    HitRec sh;
                    n_sh++;
                    if (!trace_ray<W, true>(S, sr.o, sr.d, sr.tmin, sr.tmax, sr.flags, sh, stk)) Lp = Lp + sr.contrib;
                }
            }
            if (finish) {
                Lout[slot] = make_float4(Lp.x, Lp.y, Lp.z, 0.f);
                active = false;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        n_ext += __shfl_xor(n_ext, off);
        n_sh += __shfl_xor(n_sh, off);
    }
    if (lane_id() == 0) {
        atomicAdd(&stats[0], n_ext);
        atomicAdd(&stats[1], n_sh);
    }
}

// Debug / parity entry points ------------------------------------------------
template <int W>
__global__ void __launch_bounds__(256) k_trace_rays(const SceneGPU* __restrict__ scene, const float4* __restrict__ rays, uint32_t n,
                                                    uint32_t flags, float4* __restrict__ out_hit, uint32_t* __restrict__ out_ids,
                                                    uint32_t* __restrict__ spill) {
    __shared__ uint32_t stack[kLdsStack * 256];
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const LaneStack stk{to_lds(&stack[threadIdx.x]), 256, to_global(spill + i), gridDim.x * 256};
    if (i >= n) return;
    const SceneGPU& S = *scene;
    float4 o = rays[2 * i], d = rays[2 * i + 1];
    HitRec h;
    if (flags & HH_TRACE_TERMINATE_BIT)
        trace_ray<W, true>(S, ld3(o), ld3(d), o.w, d.w, flags, h, stk);
    else
        trace_ray<W, false>(S, ld3(o), ld3(d), o.w, d.w, flags, h, stk);
    out_hit[i] = make_float4(h.found ? h.t : -1.0f, h.u, h.v, __uint_as_float(h.found ? h.inst : 0xFFFFFFFFu));
    uint32_t geom = 0xFFFFFFFFu, prim = 0xFFFFFFFFu;
    if (h.found) {
        prim = __float_as_uint(S.tris[3 * (size_t)h.tri].w);
        geom = __float_as_uint(S.tris[3 * (size_t)h.tri + 1].w) & 0x7FFFFFFFu;
    }
    out_ids[2 * i] = geom;
    out_ids[2 * i + 1] = prim;
}

template <int W>
__global__ void __launch_bounds__(256) k_primary(const SceneGPU* __restrict__ scene, const ViewGPU* __restrict__ view, uint32_t w,
                                                 uint32_t h_, uint32_t frame, float4* __restrict__ out_hit,
                                                 uint32_t* __restrict__ out_ids, uint32_t* __restrict__ spill) {
    __shared__ uint32_t stack[kLdsStack * 256];
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const LaneStack stk{to_lds(&stack[threadIdx.x]), 256, to_global(spill + i), gridDim.x * 256};
    if (i >= w * h_) return;
    uint32_t x = i % w, y = i / w;
    const SceneGPU& S = *scene;
    RNG rng = rng_init(x, y, frame);
    f3 o, d;
    generate_ray(*view, x, y, rng, o, d);
    HitRec h;
    trace_ray<W, false>(S, o, d, 0.001f, 10000.0f, 0u, h, stk);
    out_hit[i] = make_float4(h.found ? h.t : -1.0f, h.u, h.v, __uint_as_float(h.found ? h.inst : 0xFFFFFFFFu));
    uint32_t geom = 0xFFFFFFFFu, prim = 0xFFFFFFFFu;
    if (h.found) {
        prim = __float_as_uint(S.tris[3 * (size_t)h.tri].w);
        geom = __float_as_uint(S.tris[3 * (size_t)h.tri + 1].w) & 0x7FFFFFFFu;
    }
    out_ids[2 * i] = geom;
    out_ids[2 * i + 1] = prim;
}

// Output-buffer views (debug_visualization.vert:49-60, .frag:60-163). The
// reference rasterises every submesh; a fragment's perspective-correct
// attributes are the barycentric interpolation at the point where the ray
// through the fragment centre meets the triangle, so one un-jittered ray per
// pixel reproduces the pass: depth test LESS_OR_EQUAL = closest hit, cull
// none, no any-hit (the pass draws alpha-tested geometry solid), and the clip
// volume 0 <= z <= w of glm's GL-style projection = view depth in
// [2fn/(f+n), f] (near_d, far_d).
__device__ inline float4 output_buffer_value(const SceneGPU& S, uint32_t inst, uint32_t tri, float u, float v, uint32_t which) {
    const float4* tw = S.tris + 3 * (size_t)tri;
    uint32_t prim = __float_as_uint(tw[0].w);
    uint32_t geom = __float_as_uint(tw[1].w) & 0x7FFFFFFFu;
    const InstanceData& I = S.inst[inst];
    uint32_t mesh = S.inst_trav[inst].mesh;
    uint2 info = S.submesh_info[S.submesh_off[inst] + geom];
    const MaterialData& m = S.mats[info.y];
    const uint32_t* ix = S.idx + S.ibase[mesh] + 3 * (prim + info.x);
    const float4* vb = S.vtx + 5 * (size_t)S.vbase[mesh];
    const float4* Vk[3] = {vb + 5 * (size_t)ix[0], vb + 5 * (size_t)ix[1], vb + 5 * (size_t)ix[2]};
    const float b[3] = {1.0f - u - v, u, v};
    float uv0 = 0.f, uv1 = 0.f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        uv0 += Vk[k][1].x * b[k];
        uv1 += Vk[k][1].y * b[k];
    }
    f3 c = mk(0.f, 0.f, 0.f);
    if (which == 0) {  // fetch_albedo: material.albedo.rgb (Q10-converted) or texture rgb
        if (m.tex0[0] == -1) c = mk(m.albedo[0], m.albedo[1], m.albedo[2]);
        else c = ld3(sample_tex(S, m.tex0[0], uv0, uv1));
    } else if (which == 1) {
        // vertex stage: normalize(mat3(normal_matrix) * N) per vertex, then
        // interpolated without renormalising (fetch_normal without a map)
        f3 n = mk(0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 3; k++) n = n + normalize(m3_mul(I.normal, ld3(Vk[k][2]))) * b[k];
        if (m.tex0[1] != -1) {  // get_normal_from_map
            f3 tg = mk(0.f, 0.f, 0.f), bt = mk(0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                tg = tg + normalize(m3_mul(I.normal, ld3(Vk[k][3]))) * b[k];
                bt = bt + normalize(m3_mul(I.normal, ld3(Vk[k][4]))) * b[k];
            }
            f3 T = normalize(tg), B = normalize(bt), N = normalize(n);
            float4 t = sample_tex(S, m.tex0[1], uv0, uv1);
            f3 tn = normalize(mk(t.x * 2.0f - 1.0f, t.y * 2.0f - 1.0f, t.z * 2.0f - 1.0f));
            n = normalize(T * tn.x + B * tn.y + N * tn.z);
        }
        c = n * 0.5f + mk(0.5f, 0.5f, 0.5f);
    } else if (which == 2 || which == 3) {  // fetch_roughness / fetch_metallic (no MIN_ROUGHNESS clamp here)
        const int k = which - 2;
        float r;
        if (m.tex0[2 + k] == -1) r = m.roughness_metallic[k];
        else {
            float4 t = sample_tex(S, m.tex0[2 + k], uv0, uv1);
            const int ch = m.tex1[2 + k] & 3;
            r = ch == 0 ? t.x : (ch == 1 ? t.y : (ch == 2 ? t.z : t.w));
        }
        c = mk(r, r, r);
    } else {  // fetch_emissive
        if (m.tex1[0] == -1) c = mk(m.emissive[0], m.emissive[1], m.emissive[2]);
        else c = ld3(sample_tex(S, m.tex1[0], uv0, uv1));
    }
    return make_float4(c.x, c.y, c.z, 1.0f);
}

template <int W>
__global__ void __launch_bounds__(256) k_output_buffer(const SceneGPU* __restrict__ scene, const ViewGPU* __restrict__ view,
                                                       uint32_t w, uint32_t h_, float near_d, float far_d, uint32_t which,
                                                       float4* __restrict__ out, uint32_t* __restrict__ spill) {
    __shared__ uint32_t stack[kLdsStack * 256];
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const LaneStack stk{to_lds(&stack[threadIdx.x]), 256, to_global(spill + i), gridDim.x * 256};
    if (i >= w * h_) return;
    const uint32_t x = i % w, y = i / w;
    const ViewGPU& V = *view;
    const SceneGPU& S = *scene;
    const float nx = (((float)x + 0.5f) / (float)w) * 2.0f - 1.0f;
    const float ny = (((float)y + 0.5f) / (float)h_) * 2.0f - 1.0f;
    float tg[4];
#pragma unroll
    for (int r = 0; r < 4; r++) tg[r] = (V.vpi[0 * 4 + r] * nx + V.vpi[1 * 4 + r] * ny) + V.vpi[3 * 4 + r];
    const f3 cam = mk(V.cam[0], V.cam[1], V.cam[2]);
    const f3 d = normalize(mk(tg[0] / tg[3], tg[1] / tg[3], tg[2] / tg[3]) - cam);
    // the focal plane's normal is -forward
    const float cosf_ = -(d.x * V.fplane[0] + d.y * V.fplane[1] + d.z * V.fplane[2]);
    float4 res = make_float4(0.f, 0.f, 0.f, 1.f);  // clear colour (renderer.cpp:282-287)
    if (cosf_ > 0.0f) {
        HitRec hr;
        trace_ray<W, false>(S, cam, d, near_d / cosf_, far_d / cosf_, HH_TRACE_OPAQUE_BIT, hr, stk);
        if (hr.found) res = output_buffer_value(S, hr.inst, hr.tri, hr.u, hr.v, which);
    }
    out[i] = res;
}

// Ray-debug path variant (path_trace_debug.rgen, RAY_DEBUG_VIEW in
// path_trace_rgen.glsl:190-195, path_trace_rchit.glsl:500-514,548-567 and
// path_trace_rmiss.glsl:40-58): one path per debug ray through one pixel,
// launch id (i, 0) and seed num_frames; a colour from three draws before the
// camera ray; no Russian roulette (and no RR draw); every non-primary ray
// appends the segment origin -> hit (or origin + dir * tmax on a miss) as two
// DebugRayVertex records. Shadow rays only set visibility, which no segment
// depends on, so they are not traced; their draws are still consumed.
// vtx holds cap vertices; args[0] counts every appended vertex, also those
// that did not fit.
template <int W>
__global__ void __launch_bounds__(64) k_ray_debug(const SceneGPU* __restrict__ scene, const ViewGPU* __restrict__ view,
                                                  uint32_t n, uint32_t px, uint32_t py, uint32_t frame,
                                                  float4* __restrict__ vtx, uint32_t cap, uint32_t* __restrict__ args,
                                                  uint32_t* __restrict__ spill) {
    __shared__ uint32_t stack[kLdsStack * 64];
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const LaneStack stk{to_lds(&stack[threadIdx.x]), 64, to_global(spill + i), gridDim.x * 64};
    if (i >= n) return;
    const ViewGPU& V = *view;
    const SceneGPU& S = *scene;
    RNG rng = rng_init(i, 0u, frame);
    const float cr = next_float(rng) * 0.5f + 0.5f;
    const float cg = next_float(rng) * 0.5f + 0.5f;
    const float cb = next_float(rng) * 0.5f + 0.5f;
    f3 o, d;
    generate_ray(V, px, py, rng, o, d);
    f3 T = mk(1.f, 1.f, 1.f);
    float tmin = 0.001f;
    uint32_t rflags = 0;
    for (uint32_t depth = 0;; depth++) {
        HitRec h;
        trace_ray<W, false>(S, o, d, tmin, 10000.0f, rflags, h, stk);
        if (depth > 0) {
            const f3 e = o + d * (h.found ? h.t : 10000.0f);
            const uint32_t k = atomicAdd(&args[0], 2u);
            if (k + 2u <= cap) {
                vtx[2 * k + 0] = make_float4(o.x, o.y, o.z, 1.0f);
                vtx[2 * k + 1] = make_float4(cr, cg, cb, 1.0f);
                vtx[2 * k + 2] = make_float4(e.x, e.y, e.z, 1.0f);
                vtx[2 * k + 3] = make_float4(cr, cg, cb, 1.0f);
            }
        }
        if (!h.found) break;
        Surface p;
        populate(S, h.inst, h.tri, h.u, h.v, p);
        (void)direct_lighting(S, p, d, T, depth, V.shadow_bias, V.flags, rng);
        if (depth + 1 >= V.max_bounces) break;
        f3 Wi;
        float pdf;
        f3 brdf = sample_uber(p, -d, rng, Wi, pdf);
        float ct = clampf(dot(p.normal, Wi), 0.0f, 1.0f);
        T = (T * (brdf * ct)) / pdf;
        o = p.pos;
        d = Wi;
        tmin = 0.0001f;
        rflags = HH_TRACE_OPAQUE_BIT;
    }
}

// tone_map.frag:20-51
__global__ void k_tonemap(const float4* __restrict__ src, uint32_t n, float exposure, uint32_t op, uchar4* __restrict__ dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = src[i];
    float v[3] = {c.x * exposure, c.y * exposure, c.z * exposure};
    unsigned char o[3];
    for (int k = 0; k < 3; k++) {
        float x = v[k];
        if (op == 0u) x = fminf(fmaxf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f), 1.0f);
        else if (op == 1u) x = x / (1.0f + x);
        x = powf(x, 1.0f / 2.2f);
        x = fminf(fmaxf(x, 0.0f), 1.0f);
        o[k] = (unsigned char)floorf(x * 255.0f + 0.5f);
    }
    dst[i] = make_uchar4(o[0], o[1], o[2], 255);
}

// procedural_sky.frag:48-75 for one unit direction.
__device__ inline float4 hosek_eval(f3 d, const float* __restrict__ coef, float sx, float sy, float sz) {
    float cos_theta = clampf(d.y, 0.0f, 1.0f);
    float cos_gamma = clampf(d.x * sx + d.y * sy + d.z * sz, 0.0f, 1.0f);
    float gamma = acosf(cos_gamma);
    float out[3];
    for (int c = 0; c < 3; c++) {
        float A = coef[0 + c], B = coef[3 + c], C = coef[6 + c], D = coef[9 + c], E = coef[12 + c], F = coef[15 + c],
              G = coef[18 + c], H = coef[21 + c], I = coef[24 + c], Z = coef[27 + c];
        float chi = (1.0f + cos_gamma * cos_gamma) / powf(1.0f + H * H - 2.0f * cos_gamma * H, 1.5f);
        float hw = (1.0f + A * expf(B / (cos_theta + 0.01f))) *
                   (C + D * expf(E * gamma) + F * (cos_gamma * cos_gamma) + G * chi + I * sqrtf(cos_theta));
        out[c] = Z * hw;
    }
    return make_float4(out[0], out[1], out[2], 1.0f);
}

// Extension: the sky baked into an equirectangular RGBA16F map (texel
// (i, j): theta = pi (j + 0.5) / H, phi = 2 pi ((i + 0.5) / W - 0.5)).
__global__ void k_sky_bake_equirect(uint16_t* __restrict__ texels, uint32_t w, const float* __restrict__ coef, float sx,
                                    float sy, float sz) {
    const uint32_t h = w / 2;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    const float theta = HH_PI * (((float)y + 0.5f) / (float)h);
    const float phi = 2.0f * HH_PI * (((float)x + 0.5f) / (float)w - 0.5f);
    const f3 d = mk(sinf(theta) * sinf(phi), cosf(theta), -sinf(theta) * cosf(phi));
    const float4 c = hosek_eval(d, coef, sx, sy, sz);
    texels[4 * (size_t)i + 0] = __builtin_bit_cast(uint16_t, (_Float16)c.x);
    texels[4 * (size_t)i + 1] = __builtin_bit_cast(uint16_t, (_Float16)c.y);
    texels[4 * (size_t)i + 2] = __builtin_bit_cast(uint16_t, (_Float16)c.z);
    texels[4 * (size_t)i + 3] = __builtin_bit_cast(uint16_t, (_Float16)1.0f);
}

// procedural_sky.frag:48-75 evaluated at every cube texel centre.
__global__ void k_sky_bake(float4* __restrict__ faces, uint32_t n, const float* __restrict__ coef, float sx, float sy, float sz) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t total = 6u * n * n;
    if (i >= total) return;
    uint32_t face = i / (n * n), rem = i % (n * n), j = rem / n, k = rem % n;
    float sc = 2.0f * (((float)k + 0.5f) / (float)n) - 1.0f;
    float tc = 2.0f * (((float)j + 0.5f) / (float)n) - 1.0f;
    f3 d;
    switch (face) {
        case 0: d = mk(1.0f, -tc, -sc); break;
        case 1: d = mk(-1.0f, -tc, sc); break;
        case 2: d = mk(sc, 1.0f, tc); break;
        case 3: d = mk(sc, -1.0f, -tc); break;
        case 4: d = mk(sc, -tc, 1.0f); break;
        default: d = mk(-sc, -tc, -1.0f); break;
    }
    d = normalize(d);
    faces[i] = hosek_eval(d, coef, sx, sy, sz);
}

__global__ void k_fill_f4(float4* p, float4 v, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Plain streaming copy of known size: the reference point for HBM byte
// counters (FETCH_SIZE / WRITE_SIZE) and for the achievable copy bandwidth.
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

hipError_t launch_copy_probe(hipStream_t st, int num_cu, const float4* src, float4* dst, size_t n) {
    k_copy_probe<<<num_cu * 16, 256, 0, st>>>(src, dst, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Brackets one launch with HIP events on the launch stream when profiling.
struct ProfScope {
    Prof* p;
    int kind;
    hipStream_t st;
    hipEvent_t a = nullptr;
    int snap = -1;
    ProfScope(Prof* p_, int k, hipStream_t s) : p(p_), kind(k), st(s) {
        if (p && p->timing) {
            if (p->d_snaps && p->stats && p->snaps_used < Prof::kSnapCap) {
                snap = p->snaps_used++;
                (void)hipMemcpyAsync(p->d_snaps + 4 * snap, p->stats, 16, hipMemcpyDeviceToDevice, st);
            }
            a = p->ev();
            if (a) (void)hipEventRecord(a, st);
        }
    }
    ~ProfScope() {
        if (a) {
            hipEvent_t b = p->ev();
            if (b) {
                (void)hipEventRecord(b, st);
                p->recs.push_back({kind, a, b, snap});
            }
            if (snap >= 0) (void)hipMemcpyAsync(p->d_snaps + 4 * snap + 2, p->stats, 16, hipMemcpyDeviceToDevice, st);
        }
    }
};

// Queue counters: see ctr_index (extension queue lengths ping-pong between
// kCtrExt0 and kCtrExt1; each kind has kQShards segment counters).
template <int W, int BS, int MODE, bool COUNT>
static hipError_t frame_impl(const LaunchCfg& cfg, const SceneGPU* d_scenes, const ViewGPU* d_views, const ViewGPU* h_views,
                             uint32_t nv, uint32_t total, uint32_t max_bounces, Wavefront& wf, bool mega) {
    hipStream_t st = cfg.stream;
    Prof* pf = cfg.prof;
    unsigned long long* tc = pf ? pf->d_counts : nullptr;
    if (pf) pf->stats = wf.stats;
    const uint32_t grid = frame_grid(total, cfg.num_cu, cfg.tune.blocks_per_cu, BS);
    const uint32_t refill = cfg.tune.refill_lanes;
    // A wave of camera rays holds jittered samples of one or two pixels, which
    // traverse the same nodes and finish together: refilling it only when all
    // of its lanes are idle keeps it a coherent packet (measured 0.60 -> 0.55 ms
    // of extension per 1080p frame). Shadow and bounce rays diverge, and the
    // tuning's threshold is faster for them.
    const uint32_t refill_camera = 64u;
    // The continuation megakernel takes new paths once a quarter of its lanes
    // are idle (one 1080p frame: 3.21 ms at 32 lanes, 3.16 ms at 16).
    const uint32_t refill_paths = 16u;
    uint32_t* ctr = wf.counters;
    const uint64_t seg64 = queue_segment(total, grid, BS);
    if (seg64 * kQShards > wf.qcap || seg64 * kQShards > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const uint32_t seg = (uint32_t)seg64;
    // camera rays in packets when all views share one scene (the counting
    // variant keeps the per-lane traversal, whose fetches it counts)
    bool packets = !COUNT, one_scene = true;
    for (uint32_t i = 0; i < nv; i++) {
        one_scene = one_scene && h_views[i].scene == h_views[0].scene;
        packets = packets && h_views[i].scene == h_views[0].scene && h_views[i].spp >= kPacketMinSpp;
    }
    // the traversal kernels read a view only for its scene: with one scene
    // they take view 0 and skip the per-ray search of the view table
    const uint32_t nv_trace = one_scene ? 1u : nv;
    // Shadow launches on the side stream, beside the next extension launch
    // (neither reads what the other writes); `join` makes the main stream
    // wait for the last one before the next reset, shade, continuation or
    // accumulation. Profiled frames stay on one stream.
    const bool side = pf == nullptr && cfg.side != nullptr && wf.spill_threads >= 2ull * grid * BS;
    uint32_t* const spill_side = side ? wf.spill + (size_t)grid * BS * kSpillStackDepth : wf.spill;
    bool joined = true;
    hipError_t sync_err = hipSuccess;   // first failure of an event record or wait
    auto keep = [&](hipError_t e) {
        if (sync_err == hipSuccess) sync_err = e;
    };
    auto join = [&]() {
        if (!joined) keep(hipStreamWaitEvent(st, cfg.join, 0));
        joined = true;
    };
    // one bounce of every path in the queue: extension rays, shading, shadow rays
    auto bounce = [&](uint32_t d) {
        const int a = d & 1, b = a ^ 1;
        if (d == 0 && packets) {
            ProfScope ps(pf, PK_EXTEND, st);
            k_extend_packet<W, BS><<<grid, BS, 0, st>>>(d_scenes, d_views, wf.q_o[0], wf.q_d[0], ctr + ctr_index(kCtrExt0, 0),
                                                        ctr + ctr_index(kCtrExtHead, 0), seg, 0u, wf.hit, wf.hit_tri,
                                                        wf.stats);
        } else {
            ProfScope ps(pf, PK_EXTEND, st);
            k_extend<W, BS, MODE, COUNT><<<grid, BS, 0, st>>>(d_scenes, d_views, nv_trace, wf.q_o[a], wf.q_d[a], wf.q_slot[a],
                                                              ctr + ctr_index(kCtrExt0 + a, 0),
                                                              ctr + ctr_index(kCtrExtHead, 0), seg,
                                                              d == 0 ? refill_camera : refill,
                                                              d == 0 ? 0u : (uint32_t)HH_TRACE_OPAQUE_BIT, wf.hit,
                                                              wf.hit_tri, wf.stats, tc, wf.spill);
        }
        join();
        k_reset_queues<<<1, 64, 0, st>>>(ctr, kCtrExt0 + b);
        {
            ProfScope ps(pf, PK_SHADE, st);
            k_shade<BS><<<grid, BS, 0, st>>>(d_scenes, d_views, nv, d, wf.q_o[a], wf.q_d[a], wf.q_slot[a],
                                             ctr + ctr_index(kCtrExt0 + a, 0), seg,
                                             wf.hit, wf.hit_tri, wf.T, wf.L, wf.rng, wf.q_o[b], wf.q_d[b], wf.q_slot[b],
                                             ctr + ctr_index(kCtrExt0 + b, 0), wf.s_o, wf.s_d, wf.s_c, wf.s_slot,
                                             ctr + ctr_index(kCtrShadow, 0), wf.feat, wf.capacity);
        }
        hipStream_t sst = st;
        if (side) {
            keep(hipEventRecord(cfg.fork, st));
            keep(hipStreamWaitEvent(cfg.side, cfg.fork, 0));
            sst = cfg.side;
        }
        {
            ProfScope ps(pf, PK_SHADOW, sst);
            k_shadow<W, BS, MODE, COUNT><<<grid, BS, 0, sst>>>(d_scenes, d_views, nv_trace, wf.s_o, wf.s_d, wf.s_c, wf.s_slot,
                                                               ctr + ctr_index(kCtrShadow, 0),
                                                               ctr + ctr_index(kCtrShadowHead, 0), seg, refill, wf.L,
                                                               wf.stats, tc, side ? spill_side : wf.spill);
        }
        if (side) {
            keep(hipEventRecord(cfg.join, cfg.side));
            joined = false;
        }
    };
    {
        ProfScope ps(pf, PK_GENERATE, st);
        k_generate<BS><<<grid, BS, 0, st>>>(d_views, nv, total, wf.q_o[0], wf.q_d[0], wf.q_slot[0], wf.L, wf.rng, ctr, seg,
                                            max_bounces == 0 ? 1u : 0u);
    }
    // The first `cont` bounces as wavefront launches (large coherent queues),
    // then the remaining bounces of every path in one continuation launch, so
    // a pass is not held by the latency tails of the small deep-bounce
    // launches. cont = 0: every bounce as wavefront launches.
    // (the counting variant traces every bounce as wavefront launches: it
    // counts the fetches of the per-lane traversal kernels)
    const uint32_t cont = COUNT ? 0u : (mega ? 1u : cfg.tune.continuation_depth);
    if (cont > 0 && cont < max_bounces) {
        for (uint32_t d = 0; d < cont; d++) bounce(d);
        join();
        const int q = (int)(cont & 1u);  // bounce(cont - 1) queued its extension rays there
        ProfScope ps(pf, PK_MEGAKERNEL, st);
        k_megakernel<W, BS><<<grid, BS, 0, st>>>(d_scenes, d_views, nv, wf.q_o[q], wf.q_d[q], wf.q_slot[q],
                                                 ctr + ctr_index(kCtrExt0 + q, 0), ctr + ctr_index(kCtrExtHead, 0), seg,
                                                 refill_paths, wf.T, wf.rng, wf.L, wf.stats, wf.spill, cont);
    } else {
        for (uint32_t d = 0; d < max_bounces; d++) bounce(d);
        join();
    }
    uint32_t maxpix = 0;
    for (uint32_t i = 0; i < nv; i++) maxpix = std::max(maxpix, h_views[i].tile_w * h_views[i].tile_h);
    dim3 ag(std::max(1u, std::min(cdiv(maxpix, kAccBlock), 65535u)), nv);
    {
        ProfScope ps(pf, PK_ACCUMULATE, st);
        k_accumulate<<<ag, kAccBlock, 0, st>>>(d_views, nv, wf.L, wf.feat, wf.capacity);
    }
    if (sync_err != hipSuccess) return sync_err;
    return hipGetLastError();
}

hipError_t launch_frame(const LaunchCfg& cfg, const SceneGPU* d_scenes, const SceneGPU* h_scenes, const ViewGPU* d_views,
                        const ViewGPU* h_views, uint32_t nv, uint32_t total, uint32_t max_bounces, Wavefront& wf,
                        bool mega) {
    const int L = h_scenes[0].bvh_width;  // node layout id
    const int bs = cfg.tune.extend_block;
    const bool ifif = cfg.tune.traversal == kIfIf;
#define HH_FRAME(LL, BB, MM, CC) return frame_impl<LL, BB, MM, CC>(cfg, d_scenes, d_views, h_views, nv, total, max_bounces, wf, mega)
    if (cfg.prof && cfg.prof->count && cfg.prof->d_counts) {
        if (L == kLayoutB8) HH_FRAME(kLayoutB8, 256, kWhileWhile, true);
        if (L == kLayoutQ4) HH_FRAME(kLayoutQ4, 256, kWhileWhile, true);
        HH_FRAME(kLayoutB4, 256, kWhileWhile, true);
    }
#define HH_MODES(LL, BB)                               \
    do {                                               \
        if (ifif) HH_FRAME(LL, BB, kIfIf, false);      \
        HH_FRAME(LL, BB, kWhileWhile, false);          \
    } while (0)
#define HH_BLOCKS(LL)                       \
    do {                                    \
        if (bs == 64) HH_MODES(LL, 64);     \
        if (bs == 128) HH_MODES(LL, 128);   \
        HH_MODES(LL, 256);                  \
    } while (0)
    if (L == kLayoutB8) HH_BLOCKS(kLayoutB8);
    if (L == kLayoutQ4) HH_BLOCKS(kLayoutQ4);
    HH_BLOCKS(kLayoutB4);
#undef HH_BLOCKS
#undef HH_MODES
#undef HH_FRAME
}

hipError_t launch_trace_rays(const LaunchCfg& cfg, const SceneGPU* d_scene, const SceneGPU& h_scene, const float4* d_rays,
                             uint32_t n, uint32_t flags, float4* d_hit, uint32_t* d_ids, uint32_t* spill) {
    if (n == 0) return hipSuccess;
    const int L = h_scene.bvh_width;
    if (L == kLayoutB8) k_trace_rays<kLayoutB8><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_rays, n, flags, d_hit, d_ids, spill);
    else if (L == kLayoutQ4) k_trace_rays<kLayoutQ4><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_rays, n, flags, d_hit, d_ids, spill);
    else k_trace_rays<kLayoutB4><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_rays, n, flags, d_hit, d_ids, spill);
    return hipGetLastError();
}

// One thread per hit: the 24-byte record {t, u, v, instance, geometry,
// primitive} as six words.
__global__ void __launch_bounds__(256) k_pack_hits(const float4* __restrict__ hit, const uint32_t* __restrict__ ids,
                                                   uint32_t n, uint32_t* __restrict__ out) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 h = hit[i];
    const bool miss = h.x < 0.0f;
    uint32_t* o = out + 6 * (size_t)i;
    o[0] = __float_as_uint(h.x);
    o[1] = miss ? 0u : __float_as_uint(h.y);
    o[2] = miss ? 0u : __float_as_uint(h.z);
    o[3] = miss ? 0xFFFFFFFFu : __float_as_uint(h.w);
    o[4] = miss ? 0xFFFFFFFFu : ids[2 * i];
    o[5] = miss ? 0xFFFFFFFFu : ids[2 * i + 1];
}

hipError_t launch_pack_hits(hipStream_t stream, const float4* d_hit, const uint32_t* d_ids, uint32_t n, void* d_out) {
    if (n == 0) return hipSuccess;
    k_pack_hits<<<cdiv(n, 256), 256, 0, stream>>>(d_hit, d_ids, n, (uint32_t*)d_out);
    return hipGetLastError();
}

hipError_t launch_primary_hits(const LaunchCfg& cfg, const SceneGPU* d_scene, const SceneGPU& h_scene, const ViewGPU* d_view,
                               uint32_t w, uint32_t h, uint32_t frame, float4* d_hit, uint32_t* d_ids, uint32_t* spill) {
    uint32_t n = w * h;
    const int L = h_scene.bvh_width;
    if (L == kLayoutB8) k_primary<kLayoutB8><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_view, w, h, frame, d_hit, d_ids, spill);
    else if (L == kLayoutQ4) k_primary<kLayoutQ4><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_view, w, h, frame, d_hit, d_ids, spill);
    else k_primary<kLayoutB4><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_view, w, h, frame, d_hit, d_ids, spill);
    return hipGetLastError();
}

hipError_t launch_output_buffer(const LaunchCfg& cfg, const SceneGPU* d_scene, const SceneGPU& h_scene, const ViewGPU* d_view,
                                uint32_t w, uint32_t h, float near_d, float far_d, uint32_t which, float4* out,
                                uint32_t* spill) {
    const uint32_t n = w * h;
    const int L = h_scene.bvh_width;
    if (L == kLayoutB8) k_output_buffer<kLayoutB8><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_view, w, h, near_d, far_d, which, out, spill);
    else if (L == kLayoutQ4) k_output_buffer<kLayoutQ4><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_view, w, h, near_d, far_d, which, out, spill);
    else k_output_buffer<kLayoutB4><<<cdiv(n, 256), 256, 0, cfg.stream>>>(d_scene, d_view, w, h, near_d, far_d, which, out, spill);
    return hipGetLastError();
}

hipError_t launch_ray_debug(const LaunchCfg& cfg, const SceneGPU* d_scene, const SceneGPU& h_scene, const ViewGPU* d_view,
                            uint32_t n, uint32_t px, uint32_t py, uint32_t frame, float4* vtx, uint32_t cap, uint32_t* args,
                            uint32_t* spill) {
    if (n == 0) return hipSuccess;
    const int L = h_scene.bvh_width;
    if (L == kLayoutB8) k_ray_debug<kLayoutB8><<<cdiv(n, 64), 64, 0, cfg.stream>>>(d_scene, d_view, n, px, py, frame, vtx, cap, args, spill);
    else if (L == kLayoutQ4) k_ray_debug<kLayoutQ4><<<cdiv(n, 64), 64, 0, cfg.stream>>>(d_scene, d_view, n, px, py, frame, vtx, cap, args, spill);
    else k_ray_debug<kLayoutB4><<<cdiv(n, 64), 64, 0, cfg.stream>>>(d_scene, d_view, n, px, py, frame, vtx, cap, args, spill);
    return hipGetLastError();
}

hipError_t launch_tonemap(hipStream_t st,const float4* src, uint32_t n, float exposure, uint32_t op, uchar4* dst) {
    k_tonemap<<<cdiv(n, 256), 256, 0, st>>>(src, n, exposure, op, dst);
    return hipGetLastError();
}

hipError_t launch_sky_bake(hipStream_t st, float4* faces, uint32_t size, const float* d_coef, const float sun[3]) {
    uint32_t total = 6u * size * size;
    k_sky_bake<<<cdiv(total, 256), 256, 0, st>>>(faces, size, d_coef, sun[0], sun[1], sun[2]);
    return hipGetLastError();
}

hipError_t launch_sky_bake_equirect(hipStream_t st, uint16_t* texels, uint32_t width, const float* d_coef, const float sun[3]) {
    const uint32_t total = width * (width / 2);
    k_sky_bake_equirect<<<cdiv(total, 256), 256, 0, st>>>(texels, width, d_coef, sun[0], sun[1], sun[2]);
    return hipGetLastError();
}

// Counts non-finite components of a float4 image (one atomic per wave).
__global__ void k_count_nonfinite(const float4* __restrict__ p, size_t n, unsigned long long* __restrict__ out) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 v = p[i];
        c += !isfinite(v.x) + !isfinite(v.y) + !isfinite(v.z) + !isfinite(v.w);
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63u) == 0) atomicAdd(out, c);
}

hipError_t launch_count_nonfinite(hipStream_t st, const float4* p, size_t n, unsigned long long* out) {
    const uint32_t blocks = (uint32_t)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 2048));
    k_count_nonfinite<<<blocks, 256, 0, st>>>(p, n, out);
    return hipGetLastError();
}

hipError_t launch_fill_f4(hipStream_t st, float4* p, float4 v, uint32_t n) {
    if (n == 0) return hipSuccess;
    k_fill_f4<<<cdiv(n, 256), 256, 0, st>>>(p, v, n);
    return hipGetLastError();
}

}  // namespace hh
