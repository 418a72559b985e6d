// internal.hpp — host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "gpu_scene.hpp"

#define HH_TRY(expr)                             \
    do {                                         \
        hipError_t _e = (expr);                  \
        if (_e != hipSuccess) return _e;         \
    } while (0)

namespace hh {

struct BuildScratch {
    uint32_t cap = 0;
    uint32_t *keys_in = nullptr, *keys_out = nullptr, *vals_in = nullptr, *order = nullptr;
    int2 *childs = nullptr, *range = nullptr, *jobs_a = nullptr, *jobs_b = nullptr;
    uint32_t *parent = nullptr, *flags = nullptr, *counters = nullptr, *cbounds = nullptr;
    float4 *blo = nullptr, *bhi = nullptr, *aabb_lo = nullptr, *aabb_hi = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
    // clustering (PLOC) state, allocated on first use
    uint32_t p_cap = 0;
    uint32_t *p_sorted = nullptr, *p_cl[2] = {nullptr, nullptr}, *p_nn = nullptr;
    uint32_t *p_parent = nullptr, *p_count = nullptr, *p_off = nullptr;
    unsigned long long *p_flags = nullptr, *p_scan = nullptr, *p_tail = nullptr;   // p_tail: pinned host
    int2* p_child = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    // binned SAH on the GPU, allocated on first use
    uint32_t s_cap = 0, s_nmax = 0, s_cmax = 0;
    float4* s_cen = nullptr;
    uint32_t *s_ord2 = nullptr, *s_cb = nullptr, *s_bins = nullptr, *s_chunk_left = nullptr, *s_chunk_base = nullptr;
    uint32_t *s_ctr = nullptr, *s_tail = nullptr;   // s_tail: pinned host
    int4 *s_nodes = nullptr, *s_next = nullptr, *s_small = nullptr, *s_chunks = nullptr;
    int2 *s_node_chunks = nullptr, *s_meshes = nullptr;
    uint32_t s_mesh_cap = 0;
    void* s_splits = nullptr;
    hipError_t ensure(uint32_t n);
    hipError_t ensure_ploc(uint32_t n);
    hipError_t ensure_sah(uint32_t n);
    hipError_t ensure_sah_meshes(uint32_t nm);
    void release();
    void release_ploc();
    void release_sah();
    uint32_t* vals_out_tmp(uint32_t) { return order; }
    ~BuildScratch() { release(); release_ploc(); release_sah(); }
};

hipError_t build_blas(hipStream_t st, int width, int leaf_size, uint32_t n, const float4* d_tri_src, float4* d_tri_pool,
                      const float4* d_shade_src, float4* d_shade_pool, const float4* h_tri_src,
                      uint32_t tri_base, void* d_node_pool, uint32_t node_base, BuildScratch& sc, uint32_t* num_nodes,
                      float root_box[6], bool ploc, bool sah_gpu);
hipError_t build_tlas(hipStream_t st, int width, uint32_t n, const float* h_lo, const float* h_hi, void* d_node_pool,
                      uint32_t node_base, BuildScratch& sc, uint32_t* num_nodes, bool sah, bool ploc, bool sah_gpu);
hipError_t build_blas_forest(hipStream_t st, int width, int leaf_size, const std::vector<int2>& meshes, uint32_t total,
                             const float4* d_tri_src, float4* d_tri_pool, const float4* d_shade_src, float4* d_shade_pool,
                             void* d_node_pool, uint32_t node_base, BuildScratch& sc, uint32_t* roots, uint32_t* counts,
                             float* root_boxes);
hipError_t quantize_bvh4(hipStream_t st, const void* d_src, void* d_dst, uint32_t n);

// Ray queues are split into kQShards segments, each with its own length and
// read-head counter on its own 256-B line. Same-address atomics are executed
// one at a time at the memory side (about 90 M per second on MI355X, growing
// linearly with the number of distinct counters), so one counter per queue
// would serialize the per-wave allocations and refills of a whole frame.
constexpr uint32_t kQShards = 32;
constexpr uint32_t kCtrStride = 64;  // u32 words between two counters
enum CtrKind : uint32_t { kCtrExt0 = 0, kCtrExt1 = 1, kCtrShadow = 2, kCtrExtHead = 3, kCtrShadowHead = 4, kCtrKinds = 5 };
__host__ __device__ inline uint32_t ctr_index(uint32_t kind, uint32_t shard) { return (kind * kQShards + shard) * kCtrStride; }

// Entries per queue segment for `total` paths traced by a persistent grid of
// `grid` blocks of `bs` threads. A producer wave w writes to segment w % kQShards
// and emits at most one extension and one shadow ray per queue entry it reads;
// with grid-stride consumption every wave reads at most ceil(total / stride)
// entries per lane. The initial queue is dealt out in 64-entry chunks.
// Most path slots one wavefront pass may hold. Slot and queue indices are
// 32-bit, and at ~216 B of path state per slot 2^30 slots take ~232 GB of the
// 288 GB HBM; passes are shortened (fewer frames in flight) to stay below it.
constexpr uint64_t kMaxPathSlots = 1ull << 30;

inline uint64_t queue_segment(uint32_t total, uint32_t grid, uint32_t bs) {
    const uint64_t stride = (uint64_t)grid * bs, nwaves = stride / 64;
    const uint64_t per_wave = ((total + stride - 1) / stride) * 64;
    const uint64_t shade = ((nwaves + kQShards - 1) / kQShards) * per_wave;
    const uint64_t chunks = ((uint64_t)total + 63) / 64;
    const uint64_t gen = ((chunks + kQShards - 1) / kQShards) * 64;
    return shade > gen ? shade : gen;
}
inline uint32_t frame_grid(uint32_t total, int num_cu, int blocks_per_cu, int bs) {
    const uint32_t cap = (uint32_t)num_cu * (uint32_t)blocks_per_cu;
    const uint32_t need = (total + (uint32_t)bs - 1) / (uint32_t)bs;
    return need < 1 ? 1u : (need < cap ? need : cap);
}

// Wavefront state of one render call (device buffers sized for max_paths).
struct Wavefront {
    uint32_t capacity = 0;   // path slots (T, L, rng, feat)
    uint64_t qcap = 0;       // entries of each queue (kQShards segments)
    // path state by slot
    float4* T = nullptr;
    float4* L = nullptr;
    uint2* rng = nullptr;
    // extension ray queues (ping-pong)
    float4* q_o[2] = {nullptr, nullptr};
    float4* q_d[2] = {nullptr, nullptr};
    uint32_t* q_slot[2] = {nullptr, nullptr};
    // hits of the current extension queue
    float4* hit = nullptr;       // t, u, v, bits(instance)
    uint32_t* hit_tri = nullptr; // global triangle index, 0xFFFFFFFF on miss
    // shadow queue
    float4* s_o = nullptr;
    float4* s_d = nullptr;
    float4* s_c = nullptr;       // contribution rgb, bits(ray flags)
    uint32_t* s_slot = nullptr;
    // first-hit features by slot, written at depth 0 when a view has feature
    // planes: feat[slot] = (albedo, 1/(1+t)), feat[capacity + slot] = (normal, 0)
    float4* feat = nullptr;
    // kCtrKinds x kQShards counters, see ctr_index
    uint32_t* counters = nullptr;
    unsigned long long* stats = nullptr;  // [0] extension rays, [1] shadow rays
    // traversal stack overflow, kSpillStackDepth words per launched thread
    uint32_t* spill = nullptr;
    size_t spill_threads = 0;
    hipError_t ensure(uint32_t n, uint64_t queue_entries);
    hipError_t ensure_spill(size_t threads);
    void release();
    ~Wavefront() { release(); }
};

struct Tuning {
    int bvh_width = 4;
    int extend_block = 256;
    int shade_block = 256;
    uint32_t queue_capacity = 0;
    int blocks_per_cu = 8;
    int leaf_size = 4;
    uint32_t frames_in_flight = 4;
    bool bvh_compressed = false;
    uint32_t bvh_builder = 3;   // 0: LBVH on the GPU; 1: binned SAH on the host; 2: PLOC on the GPU; 3: binned SAH on the GPU
    uint32_t refill_lanes = 24;  // idle lanes that make a traversal wave take new rays
    int traversal = 0;           // 0 while-while, 1 if-if step shape
    uint32_t continuation_depth = 0;  // bounces before the continuation launch (0 = none)
    int layout() const { return bvh_compressed && bvh_width == 4 ? kLayoutQ4 : bvh_width; }
};

// Optional per-kernel profiling of a frame: HIP events around every launch
// (timing) and the counting traversal variant (algorithmic bytes).
enum ProfKind { PK_GENERATE = 0, PK_EXTEND, PK_SHADE, PK_SHADOW, PK_ACCUMULATE, PK_MEGAKERNEL, PK_COUNT };
struct Prof {
    bool timing = false;
    bool count = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
        int snap;   // index into d_snaps, -1 = none
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    double ms[PK_COUNT] = {0};
    unsigned long long launches[PK_COUNT] = {0};
    unsigned long long* d_counts = nullptr;  // [ext nodes, ext tris, ext insts, shadow nodes, shadow tris, shadow insts]
    // Rays traced per launch kind: the frame's ray statistics (extension,
    // shadow) copied before and after every timed launch (4 words per launch,
    // kSnapCap launches between two reads), attributed to the launch's kind.
    static constexpr int kSnapCap = 4096;
    unsigned long long* d_snaps = nullptr;
    int snaps_used = 0;
    const unsigned long long* stats = nullptr;   // set by the frame being profiled
    unsigned long long rays[PK_COUNT][2] = {};
    hipEvent_t ev() {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
};

// Kernel launchers (integrator.hip)
struct LaunchCfg {
    hipStream_t stream;
    int num_cu;
    Tuning tune;
    Prof* prof = nullptr;
    // A second stream and two events: when set (and not profiling), each
    // bounce's shadow launch runs on `side` beside the next bounce's
    // extension launch, joined before anything else touches the shadow queue
    // or the radiance. Needs a spill area for twice the launched threads.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
hipError_t launch_frame(const LaunchCfg& cfg, const SceneGPU* d_scenes, const SceneGPU* h_scenes, const ViewGPU* d_views,
                        const ViewGPU* h_views, uint32_t num_views, uint32_t total_paths, uint32_t max_bounces,
                        Wavefront& wf, bool megakernel);
hipError_t launch_trace_rays(const LaunchCfg& cfg, const SceneGPU* d_scene, const SceneGPU& h_scene, const float4* d_rays,
                             uint32_t n, uint32_t flags, float4* d_hit, uint32_t* d_hit_ids, uint32_t* spill);
// Converts the raw hits of launch_trace_rays (t,u,v,bits(instance) + geometry,
// primitive) into 24-byte hh_hit records {t, u, v, instance, geometry,
// primitive} in device memory, misses as hh_trace_rays reports them
// (u = v = 0, every index 0xFFFFFFFF).
hipError_t launch_pack_hits(hipStream_t stream, const float4* d_hit, const uint32_t* d_hit_ids, uint32_t n,
                            void* d_out);
hipError_t launch_primary_hits(const LaunchCfg& cfg, const SceneGPU* d_scene, const SceneGPU& h_scene, const ViewGPU* d_view,
                               uint32_t w, uint32_t h, uint32_t frame, float4* d_hit, uint32_t* d_hit_ids, uint32_t* spill);
hipError_t launch_output_buffer(const LaunchCfg& cfg, const SceneGPU* d_scene, const SceneGPU& h_scene, const ViewGPU* d_view,
                                uint32_t w, uint32_t h, float near_d, float far_d, uint32_t which, float4* out, uint32_t* spill);
hipError_t launch_ray_debug(const LaunchCfg& cfg, const SceneGPU* d_scene, const SceneGPU& h_scene, const ViewGPU* d_view,
                            uint32_t n, uint32_t px, uint32_t py, uint32_t frame, float4* vtx, uint32_t cap, uint32_t* args,
                            uint32_t* spill);
hipError_t launch_tonemap(hipStream_t st, const float4* src, uint32_t n, float exposure, uint32_t op, uchar4* dst);
hipError_t launch_sky_bake(hipStream_t st, float4* faces, uint32_t size, const float* d_coef, const float sun[3]);
hipError_t launch_sky_bake_equirect(hipStream_t st, uint16_t* texels, uint32_t width, const float* d_coef, const float sun[3]);

// Host Hosek-Wilkie fit (sky_host.cpp)
void sky_coefficients(float turbidity, float albedo, float normalized_sun_y, const float sun_dir[3], float out[30]);
hipError_t launch_fill_f4(hipStream_t st, float4* p, float4 v, uint32_t n);
hipError_t launch_count_nonfinite(hipStream_t st, const float4* p, size_t n, unsigned long long* out);

// The learned denoiser (denoise.hip): U-Net of 14 3x3 convolutions + a 1x1
// head, BF16 MFMA kernels with FP32 accumulation, Adam on FP32 master weights.
constexpr int kDnLayers = 14;
namespace dn {
size_t net_num_params();
void net_he_init(uint64_t seed, float* p);
struct Net {
    // parameters: FP32 master vector (layer by layer: W (Cout, 3, 3, Cin), b;
    // then the head W (3, 32), b), gradients, Adam moments, BF16 copies
    size_t nparams = 0;
    float *master = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    size_t w_off[kDnLayers + 1] = {0};
    __bf16* wf[kDnLayers] = {nullptr};   // forward layout
    __bf16* wb[kDnLayers] = {nullptr};   // input-gradient layout
    __bf16* whead = nullptr;
    uint32_t steps = 0;
    // activations for n images of H x W (multiples of 8)
    uint32_t n = 0, H = 0, W = 0;
    size_t cap_px = 0;
    bool cap_train = false;
    __bf16* xin = nullptr;
    float* resid = nullptr;   // float4 per pixel: log1p radiance
    float* out = nullptr;     // float4 per pixel: network output (log1p space)
    float* dout = nullptr;
    __bf16* act[kDnLayers] = {nullptr};
    __bf16* pool[3] = {nullptr};
    __bf16* gy[2] = {nullptr, nullptr};
    float* gskip[3] = {nullptr};
    float* gpool[3] = {nullptr};
    float* loss = nullptr;

    hipError_t init();
    hipError_t zero_state(hipStream_t st);
    hipError_t refresh_weights(hipStream_t st);
    hipError_t ensure(uint32_t N, uint32_t Hp, uint32_t Wp, bool training);
    hipError_t pack(hipStream_t st, const float* raw, const float4* rad, const float4* aux, uint32_t held, uint32_t h, uint32_t w);
    hipError_t forward(hipStream_t st, float4* surface, uint32_t h, uint32_t w);
    hipError_t loss_and_backward(hipStream_t st, const float* target, uint32_t h, uint32_t w, uint32_t kind);
    hipError_t adam(hipStream_t st, float lr);
    void release_acts();
    void release();
    ~Net() { release(); }
};
}  // namespace dn
hipError_t launch_copy_probe(hipStream_t st, int num_cu, const float4* src, float4* dst, size_t n);

}  // namespace hh
