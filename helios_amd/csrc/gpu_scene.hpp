// gpu_scene.hpp — device-resident scene, BVH and wavefront-queue layouts.
//
// HBM layout of one committed scene:
//   nodes     wide BVH nodes of every BLAS and the TLAS in one pool
//             (BVH4: 128 B = one L2 line; BVH8: 256 B)
//   tris      watertight-test triangles in BLAS leaf order, 3 x float4 = 48 B:
//             v0.xyz|prim  v1.xyz|geom (bit 31: geometry opaque)  v2.xyz|instanceless 0
//   inst      per instance: world->object 3x4, BLAS root, (64 B)
//   the reference records, byte compatible: MaterialData 80 B
//   (scene.cpp:25-32), LightData 64 B (:36-42), InstanceData model/normal
//   matrices (:46-52), per-instance submesh table uvec2 (:1125-1137),
//   80-B Vertex and u32 index pools per mesh (mesh.h:10-17).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hh {

// Ray / integrator flag bits (values of HH_TRACE_* and HH_FLAG_* in helios_hip.h).
constexpr uint32_t HH_TRACE_OPAQUE_BIT = 0x1u;
constexpr uint32_t HH_TRACE_TERMINATE_BIT = 0x4u;
constexpr uint32_t HH_FLAG_NO_CLAMP_BIT = 0x1u;
constexpr uint32_t HH_FLAG_DIRECT_ONLY_BIT = 0x2u;
constexpr uint32_t HH_FLAG_MEGAKERNEL_BIT = 0x4u;
constexpr uint32_t HH_FLAG_LIGHT_POWER_BIT = 0x8u;

// Child reference encoding shared by all wide nodes:
//   bit 31 clear          -> internal node index (global pool index)
//   bit 31 set            -> leaf: bits 0..26 first triangle (global), bits
//                            27..30 count-1 (1..16 triangles); in the TLAS the
//                            "first triangle" is the instance index.
//   0xFFFFFFFF            -> stack marker "leave the BLAS, restore world ray"
// Traversal stack entries kept outside LDS, per launched thread.
constexpr int kSpillStackDepth = 104;

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMarker = 0xFFFFFFFFu;
constexpr uint32_t kLeafStartMask = 0x07FFFFFFu;
constexpr uint32_t kMaxLeafTris = 16;
__host__ __device__ inline uint32_t make_leaf(uint32_t start, uint32_t count) {
    return kLeafBit | ((count - 1u) << 27) | start;
}
__host__ __device__ inline uint32_t leaf_start(uint32_t ref) { return ref & kLeafStartMask; }
__host__ __device__ inline uint32_t leaf_count(uint32_t ref) { return ((ref >> 27) & 15u) + 1u; }

template <int W>
struct alignas(16) WideNode {
    float lox[W], hix[W], loy[W], hiy[W], loz[W], hiz[W];
    uint32_t child[W];
    uint32_t nchild;
    uint32_t pad[W - 1];
};
static_assert(sizeof(WideNode<4>) == 128, "BVH4 node must be one 128-B line");
static_assert(sizeof(WideNode<8>) == 256, "BVH8 node must be two 128-B lines");

// Quantized BVH4 node, 64 B (four 16-B loads instead of eight): child boxes
// as 8-bit offsets from the node origin in units of 2^e per axis. The decoded
// plane fmaf(q, 2^e, origin) is rounded outward at build time, so a decoded
// box always contains the exact child box and the robust slab test stays
// conservative.
struct alignas(16) QNode4 {
    float ox, oy, oz;
    int8_t ex, ey, ez;
    uint8_t nchild;
    uint8_t qlox[4], qhix[4], qloy[4], qhiy[4], qloz[4], qhiz[4];
    uint32_t child[4];
    uint32_t pad[2];
};
static_assert(sizeof(QNode4) == 64, "quantized BVH4 node is 64 B");

// Node layouts a scene's BVH can use (SceneGPU::bvh_width holds one of these).
constexpr int kLayoutB4 = 4;    // WideNode<4>, float boxes, 128 B
constexpr int kLayoutB8 = 8;    // WideNode<8>, float boxes, 256 B
constexpr int kLayoutQ4 = 16;   // QNode4, quantized boxes, 64 B
__host__ __device__ constexpr int layout_width(int layout) { return layout == kLayoutB8 ? 8 : 4; }

struct alignas(16) InstTrav {
    float w2o[12];        // row-major 3x4 world->object
    uint32_t blas_root;   // global node index
    uint32_t mesh;        // global mesh index (reference mesh_index)
    // Alpha of a candidate hit on this instance's non-opaque geometry when all
    // of it uses one alpha source: a texture slot (the albedo texture, alpha
    // channel), kAlphaConst (alpha = __uint_as_float(alpha_value)) or
    // kAlphaMixed (look the material up per geometry).
    uint32_t alpha_src;
    uint32_t alpha_value;
};
constexpr uint32_t kAlphaConst = 0xFFFFFFFEu;
constexpr uint32_t kAlphaMixed = 0xFFFFFFFFu;

// MaterialData (scene.cpp:25-32)
struct alignas(16) MaterialData {
    int32_t tex0[4];
    int32_t tex1[4];
    float albedo[4];
    float emissive[4];
    float roughness_metallic[4];
};
// LightData (scene.cpp:36-42)
struct alignas(16) LightData {
    float d0[4], d1[4], d2[4], d3[4];
};
// InstanceData (scene.cpp:46-52)
struct alignas(16) InstanceData {
    float model[16];
    float normal[16];
    uint32_t mesh_index;
    float padding[3];
};
static_assert(sizeof(InstanceData) == 144, "InstanceData is 144 B in the reference");

struct TexDesc {
    uint32_t offset;  // in float4 texels
    uint32_t w, h, pad;
};

// Everything a kernel needs to trace and shade in one scene.
struct SceneGPU {
    const void* nodes;            // WideNode<W> pool
    const float4* tris;           // 3 per triangle
    const float4* shade;          // 4 per triangle: uv0 uv1 | uv2 n0.xy | n0.z n1 | n2 (leaf order)
    const InstTrav* inst_trav;
    const InstanceData* inst;
    const MaterialData* mats;
    const LightData* lights;
    const uint32_t* submesh_off;  // per instance: offset (in uint2) into submesh_info
    const uint2* submesh_info;    // (base_index/3, material)
    const float4* vtx;            // 80-B vertices as 5 float4
    const uint32_t* vbase;        // per mesh: first vertex in vtx
    const uint32_t* idx;
    const uint32_t* ibase;        // per mesh: first index in idx
    const TexDesc* tex;
    const float4* texels;
    const float4* env;            // 6 faces, env_size^2 texels each
    uint32_t env_size;
    uint32_t num_lights;
    uint32_t tlas_root;
    uint32_t bvh_width;
    // equirectangular environment (RGBA16F, env_eq_w x env_eq_w/2); used
    // instead of the cube when env_eq_w > 0
    const uint16_t* env_eq;
    uint32_t env_eq_w;
    // power-proportional emitters: every emissive triangle of every instance
    uint32_t num_emit;
    uint32_t num_area;            // area-light records at the front of `lights`
    const float4* emit_tri;       // 3 per triangle: world v0 | area, v1, v2
    const float4* emit_le;        // emission rgb | selection probability
    const float* emit_cdf;        // inclusive cumulative probability
    // Traversal limit counters (kFault*), incremented only when a ray's
    // traversal was cut short by a stack bound or the trip bound.
    unsigned long long* faults;
};

// Indices into SceneGPU::faults.
constexpr int kFaultLaneStack = 0;    // a per-lane stack push dropped (LDS + spill full)
constexpr int kFaultPacketStack = 1;  // a wave-uniform packet stack push dropped
constexpr int kFaultTrips = 2;        // a ray stopped at the node/leaf visit bound
constexpr int kNumFaults = 3;

// Per-view launch constants (the 192-B PathTraceConsts block, reduced to the
// fields the path reads) plus the batching bookkeeping.
struct ViewGPU {
    float vpi[16];       // view_proj_inverse, column-major
    float cam[4], up[4], right[4], fplane[4];
    float aperture;
    float shadow_bias;
    uint32_t width, height;
    uint32_t tile_x, tile_y, tile_w, tile_h;  // launch region (tiled mode)
    uint32_t spp;
    uint32_t first_frame;
    uint32_t max_bounces;
    uint32_t rr_start;   // first depth at which Russian roulette may end a path
    uint32_t flags;
    uint32_t scene;      // index into the scene table
    uint32_t slot_base;  // first path slot of this view
    float4* accum;       // RGBA32F surface
    // Optional feature planes of the surface (nullptr = none): aux[p] =
    // (albedo rgb, 1/(1+t)) and aux[width*height + p] = (shading normal xyz,
    // M2 of the sample luminance), accumulated like the radiance.
    float4* aux;
};

}  // namespace hh
