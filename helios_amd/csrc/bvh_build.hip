// bvh_build.hip — GPU construction of the wide BVHs (one BLAS per mesh, one
// TLAS over mesh-node instances).
//
// Replaces the driver's acceleration-structure builds
// (src/engine/resource/mesh.cpp:46-117: one BLAS per mesh, one geometry per
// submesh; src/engine/gfx/renderer.cpp:113-181: TLAS over <= 1024 instances).
//
// Method: linear BVH over 30-bit Morton codes (Karras, "Maximizing Parallelism
// in the Construction of BVHs, Octrees, and k-d Trees", HPG 2012): sort,
// radix tree, bottom-up bounds with arrival counters; then a top-down collapse
// of the binary tree into W-wide nodes (W = 4 or 8) that repeatedly opens the
// child with the largest surface area, keeping subtrees of <= leaf_size
// primitives as leaves (each LBVH subtree covers a contiguous range of the
// sorted primitives, so a leaf is a range).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "gpu_scene.hpp"
#include "internal.hpp"

namespace hh {

namespace {

__device__ inline uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// float <-> order-preserving uint for atomic min/max
__device__ inline uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float o2f(uint32_t u) {
    uint32_t v = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    __builtin_memcpy(&f, &v, 4);
    return f;
}

__global__ void k_tri_aabbs(const float4* __restrict__ tri, uint32_t n, float4* __restrict__ lo, float4* __restrict__ hi,
                            uint32_t* __restrict__ cbounds) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float c[3] = {0, 0, 0};
    bool ok = i < n;
    if (ok) {
        float4 a = tri[3 * i], b = tri[3 * i + 1], d = tri[3 * i + 2];
        float4 l = make_float4(fminf(a.x, fminf(b.x, d.x)), fminf(a.y, fminf(b.y, d.y)), fminf(a.z, fminf(b.z, d.z)), 0.f);
        float4 h = make_float4(fmaxf(a.x, fmaxf(b.x, d.x)), fmaxf(a.y, fmaxf(b.y, d.y)), fmaxf(a.z, fmaxf(b.z, d.z)), 0.f);
        lo[i] = l;
        hi[i] = h;
        c[0] = 0.5f * (l.x + h.x);
        c[1] = 0.5f * (l.y + h.y);
        c[2] = 0.5f * (l.z + h.z);
    }
    // wave-level min/max, then one atomic per wave
    for (int k = 0; k < 3; k++) {
        float mn = ok ? c[k] : INFINITY, mx = ok ? c[k] : -INFINITY;
        for (int off = 32; off > 0; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off));
            mx = fmaxf(mx, __shfl_xor(mx, off));
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMin(&cbounds[k], f2o(mn));
            atomicMax(&cbounds[3 + k], f2o(mx));
        }
    }
}

__global__ void k_box_centroids(const float4* __restrict__ lo, const float4* __restrict__ hi, uint32_t n,
                                uint32_t* __restrict__ cbounds) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = i < n;
    float c[3] = {0, 0, 0};
    if (ok) {
        float4 l = lo[i], h = hi[i];
        c[0] = 0.5f * (l.x + h.x);
        c[1] = 0.5f * (l.y + h.y);
        c[2] = 0.5f * (l.z + h.z);
    }
    for (int k = 0; k < 3; k++) {
        float mn = ok ? c[k] : INFINITY, mx = ok ? c[k] : -INFINITY;
        for (int off = 32; off > 0; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off));
            mx = fmaxf(mx, __shfl_xor(mx, off));
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMin(&cbounds[k], f2o(mn));
            atomicMax(&cbounds[3 + k], f2o(mx));
        }
    }
}

__global__ void k_init_bounds(uint32_t* cb) {
    if (threadIdx.x < 3) {
        cb[threadIdx.x] = 0xFFFFFFFFu;
        cb[3 + threadIdx.x] = 0u;
    }
}

__global__ void k_morton(const float4* __restrict__ lo, const float4* __restrict__ hi, uint32_t n,
                         const uint32_t* __restrict__ cbounds, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 l = lo[i], h = hi[i];
    float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float mn = o2f(cbounds[k]), mx = o2f(cbounds[3 + k]);
        float ext = mx - mn;
        float t = ext > 0.f ? (c[k] - mn) / ext : 0.5f;
        q[k] = (uint32_t)fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
    }
    keys[i] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    vals[i] = i;
}

__device__ inline int delta(const uint32_t* keys, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((uint32_t)(i ^ j));
    return __clz(a ^ b);
}

// Karras 2012, Fig. 4. Internal nodes 0..n-2; leaf j has binary ref n-1+j.
__global__ void k_radix_tree(const uint32_t* __restrict__ keys, int n, int2* __restrict__ childs,
                             uint32_t* __restrict__ parent, int2* __restrict__ range) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int lo_ = min(i, j), hi_ = max(i, j);
    int left = (lo_ == gamma) ? (n - 1 + gamma) : gamma;
    int right = (hi_ == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    childs[i] = make_int2(left, right);
    range[i] = make_int2(lo_, hi_);
    parent[left] = (uint32_t)i;
    parent[right] = (uint32_t)i;
}

__device__ inline float4 aload4(const float4* p) {
    const uint32_t* q = (const uint32_t*)p;
    float4 r;
    r.x = __uint_as_float(__hip_atomic_load(q + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.y = __uint_as_float(__hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.z = __uint_as_float(__hip_atomic_load(q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.w = __uint_as_float(__hip_atomic_load(q + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    return r;
}

// Bottom-up bounds: node boxes live in blo/bhi indexed by binary ref
// (internal 0..n-2, leaves n-1..2n-2); .w of blo holds the primitive count.
__global__ void k_refit(const float4* __restrict__ plo, const float4* __restrict__ phi, const uint32_t* __restrict__ order,
                        int n, const int2* __restrict__ childs, const uint32_t* __restrict__ parent, uint32_t* __restrict__ flags,
                        float4* blo, float4* bhi) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t p = order[j];
    float4 l = plo[p], h = phi[p];
    l.w = 1.0f;
    blo[n - 1 + j] = l;
    bhi[n - 1 + j] = h;
    if (n == 1) return;
    uint32_t node = parent[n - 1 + j];
    for (;;) {
        __threadfence();
        uint32_t old = atomicAdd(&flags[node], 1u);
        if (old == 0) return;
        __threadfence();
        int2 c = childs[node];
        float4 al = aload4(&blo[c.x]), ah = aload4(&bhi[c.x]);
        float4 bl = aload4(&blo[c.y]), bh = aload4(&bhi[c.y]);
        float4 rl = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), al.w + bl.w);
        float4 rh = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.f);
        uint32_t* ql = (uint32_t*)&blo[node];
        uint32_t* qh = (uint32_t*)&bhi[node];
        __hip_atomic_store(ql + 0, __float_as_uint(rl.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ql + 1, __float_as_uint(rl.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ql + 2, __float_as_uint(rl.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ql + 3, __float_as_uint(rl.w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(qh + 0, __float_as_uint(rh.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(qh + 1, __float_as_uint(rh.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(qh + 2, __float_as_uint(rh.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(qh + 3, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (node == 0) return;
        node = parent[node];
    }
}

__device__ inline float half_area(float4 l, float4 h) {
    float dx = h.x - l.x, dy = h.y - l.y, dz = h.z - l.z;
    return dx * dy + dy * dz + dz * dx;
}

// One wide node per job: job = (binary ref, wide index).
template <int W>
__global__ void k_collapse(const int2* __restrict__ jobs, const uint32_t* __restrict__ njobs_p, int2* __restrict__ out_jobs,
                           uint32_t* __restrict__ out_njobs, uint32_t* __restrict__ node_counter, int n, int leaf_size,
                           const int2* __restrict__ childs, const int2* __restrict__ range, const float4* __restrict__ blo,
                           const float4* __restrict__ bhi, const uint32_t* __restrict__ order, WideNode<W>* __restrict__ nodes,
                           uint32_t node_base, uint32_t leaf_base, int leaf_is_prim) {
    uint32_t jid = blockIdx.x * blockDim.x + threadIdx.x;
    if (jid >= *njobs_p) return;
    int2 job = jobs[jid];
    int bref = job.x;
    auto is_internal = [&](int r) { return r < n - 1; };
    auto count_of = [&](int r) { return is_internal(r) ? (int)blo[r].w : 1; };
    auto expandable = [&](int r) { return is_internal(r) && count_of(r) > leaf_size; };
    int c[W];
    int nc = 0;
    if (!expandable(bref)) {
        c[nc++] = bref;
    } else {
        c[nc++] = childs[bref].x;
        c[nc++] = childs[bref].y;
        while (nc < W) {
            int best = -1;
            float ba = -1.f;
            for (int k = 0; k < nc; k++)
                if (expandable(c[k])) {
                    float a = half_area(blo[c[k]], bhi[c[k]]);
                    if (a > ba) { ba = a; best = k; }
                }
            if (best < 0) break;
            int r = c[best];
            c[best] = childs[r].x;
            c[nc++] = childs[r].y;
        }
    }
    // absent children get all-infinite boxes, which every ray misses (the
    // traversal's box test needs no child count)
    WideNode<W> nd;
    for (int k = 0; k < W; k++) {
        nd.lox[k] = nd.loy[k] = nd.loz[k] = INFINITY;
        nd.hix[k] = nd.hiy[k] = nd.hiz[k] = INFINITY;
        nd.child[k] = 0;
    }
    nd.nchild = (uint32_t)nc;
    for (int k = 0; k < W - 1; k++) nd.pad[k] = 0;
    for (int k = 0; k < nc; k++) {
        int r = c[k];
        float4 l = blo[r], h = bhi[r];
        nd.lox[k] = l.x; nd.loy[k] = l.y; nd.loz[k] = l.z;
        nd.hix[k] = h.x; nd.hiy[k] = h.y; nd.hiz[k] = h.z;
        if (expandable(r)) {
            uint32_t idx = atomicAdd(node_counter, 1u);
            uint32_t slot = atomicAdd(out_njobs, 1u);
            out_jobs[slot] = make_int2(r, (int)idx);
            nd.child[k] = node_base + idx;
        } else {
            int first, cnt;
            if (is_internal(r)) { first = range[r].x; cnt = range[r].y - range[r].x + 1; }
            else { first = r - (n - 1); cnt = 1; }
            nd.child[k] = leaf_is_prim ? make_leaf(order[first], 1) : make_leaf(leaf_base + (uint32_t)first, (uint32_t)cnt);
        }
    }
    nodes[job.y] = nd;
}

__global__ void k_gather_tris(const float4* __restrict__ src, const uint32_t* __restrict__ order, uint32_t n,
                              float4* __restrict__ dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = order[i];
    dst[3 * i] = src[3 * p];
    dst[3 * i + 1] = src[3 * p + 1];
    dst[3 * i + 2] = src[3 * p + 2];
}

// The per-triangle shading records follow their triangles into leaf order.
__global__ void k_gather_shade(const float4* __restrict__ src, const uint32_t* __restrict__ order, uint32_t n,
                               float4* __restrict__ dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = order[i];
#pragma unroll
    for (int k = 0; k < 4; k++) dst[4 * i + k] = src[4 * p + k];
}

// Quantization of one axis of one child box: the decoded planes
// fmaf(q, 2^e, origin) must satisfy lo' <= lo and hi' >= hi exactly in float.
__device__ inline bool quant_axis(float o, float s, float clo, float chi, uint8_t& qlo, uint8_t& qhi) {
    int a = (int)floor(((double)clo - (double)o) / (double)s);
    a = min(max(a, 0), 255);
    while (a > 0 && fmaf((float)a, s, o) > clo) a--;
    int b = (int)ceil(((double)chi - (double)o) / (double)s);
    b = min(max(b, 0), 255);
    while (b < 255 && fmaf((float)b, s, o) < chi) b++;
    qlo = (uint8_t)a;
    qhi = (uint8_t)b;
    return fmaf((float)a, s, o) <= clo && fmaf((float)b, s, o) >= chi;
}

__device__ inline float exp2i(int e) { return __int_as_float((e + 127) << 23); }

// BVH4 (float boxes, 128 B) -> quantized BVH4 (64 B), same topology.
__global__ void k_quantize4(const WideNode<4>* __restrict__ src, QNode4* __restrict__ dst, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    WideNode<4> s = src[i];
    QNode4 q;
    q.nchild = (uint8_t)s.nchild;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const float* L[3] = {s.lox, s.loy, s.loz};
    const float* H[3] = {s.hix, s.hiy, s.hiz};
    for (uint32_t k = 0; k < s.nchild; k++)
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], L[a][k]);
            hi[a] = fmaxf(hi[a], H[a][k]);
        }
    if (s.nchild == 0)
        for (int a = 0; a < 3; a++) lo[a] = hi[a] = 0.f;
    uint8_t* QL[3] = {q.qlox, q.qloy, q.qloz};
    uint8_t* QH[3] = {q.qhix, q.qhiy, q.qhiz};
    int8_t* E[3] = {&q.ex, &q.ey, &q.ez};
    float* O[3] = {&q.ox, &q.oy, &q.oz};
    for (int a = 0; a < 3; a++) {
        *O[a] = lo[a];
        double ext = (double)hi[a] - (double)lo[a];
        int e = -126;
        if (ext > 0.0) {
            int fe;
            frexp(ext / 255.0, &fe);
            e = max(fe, -126);
        }
        for (;; e++) {
            bool ok = true;
            float sc = exp2i(e);
            for (int k = 0; k < 4; k++) {
                if ((uint32_t)k < s.nchild) ok &= quant_axis(lo[a], sc, L[a][k], H[a][k], QL[a][k], QH[a][k]);
                else QL[a][k] = QH[a][k] = 0;
            }
            if (ok || e >= 127) break;
        }
        *E[a] = (int8_t)e;
    }
    for (int k = 0; k < 4; k++) q.child[k] = s.child[k];
    q.pad[0] = q.pad[1] = 0;
    dst[i] = q;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Binned SAH arithmetic shared by the host builder and the GPU builder, so
// both choose the same splits: 32 centroid bins per axis, bin of a centroid
// c in [cl, ch] = trunc((c - cl) * (32 / (ch - cl))) capped at 31, half
// surface area dx*dy + dy*dz + dz*dx, and the cost of a split = left count x
// left area + right count x right area, evaluated in this order (the library
// is built without floating-point contraction).
// ---------------------------------------------------------------------------
constexpr int kSahBins = 32;

__host__ __device__ inline int sah_bin(float c, float cl, float scale) {
    const float t = (c - cl) * scale;
    return t < float(kSahBins - 1) ? int(t) : kSahBins - 1;   // NaN -> last bin
}
__host__ __device__ inline float sah_area(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    return dx * dy + dy * dz + dz * dx;
}
__host__ __device__ inline float sah_cost(int lc, float la, int rc, float ra) {
    return float(lc) * la + float(rc) * ra;
}

// ---------------------------------------------------------------------------
// PLOC: parallel locally-ordered clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 24(3), 2018). The clusters start as the primitives in Morton
// order; every round each cluster finds its nearest neighbour among the
// kPlocRadius clusters on either side (distance = surface area of the union
// box), mutual nearest neighbours merge into a new node that takes the lower
// position, and the cluster array is compacted, until one cluster is left.
// Ties are broken by the neighbour's position, which orders the candidate
// pairs of every cluster by one total order on pairs (area, lower position,
// higher position); the smallest pair of a round is then mutual, so every
// round merges at least one pair.
//
// Node ids: leaf k = k-th primitive in Morton order (0..n-1); internal nodes
// n..2n-2 in creation order, so the root is 2n-2. A merged node's children
// are (cluster at the lower position, cluster at the higher position).
// ---------------------------------------------------------------------------
constexpr int kPlocRadius = 16;
constexpr int kPlocBlock = 256;

__global__ void k_ploc_init(const float4* __restrict__ lo, const float4* __restrict__ hi, const uint32_t* __restrict__ sorted,
                            uint32_t n, float4* __restrict__ blo, float4* __restrict__ bhi, uint32_t* __restrict__ clusters,
                            uint32_t* __restrict__ count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = sorted[i];
    float4 l = lo[p], h = hi[p];
    blo[i] = make_float4(l.x, l.y, l.z, 0.f);
    bhi[i] = make_float4(h.x, h.y, h.z, 0.f);
    clusters[i] = i;
    count[i] = 1u;
}

__device__ inline float union_half_area(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy, float cz,
                                        float dx, float dy, float dz) {
    // boxes (a lo, b hi) and (c lo, d hi)
    float ex = fmaxf(bx, dx) - fminf(ax, cx), ey = fmaxf(by, dy) - fminf(ay, cy), ez = fmaxf(bz, dz) - fminf(az, cz);
    float s = ex * ey + ey * ez + ez * ex;
    return s == s ? s : INFINITY;   // a NaN box never wins a tie-free comparison otherwise
}

// nn[i] = the position j (|i - j| <= kPlocRadius, j != i) minimizing
// (union area, j). The block stages the boxes of its positions and of the
// kPlocRadius positions on either side in LDS.
__global__ void __launch_bounds__(kPlocBlock) k_ploc_nn(const uint32_t* __restrict__ clusters, uint32_t m,
                                                        const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                        uint32_t* __restrict__ nn) {
    constexpr int S = kPlocBlock + 2 * kPlocRadius;
    __shared__ float box[6][S];
    const int base = (int)(blockIdx.x * kPlocBlock);
    for (int k = threadIdx.x; k < S; k += kPlocBlock) {
        int idx = base - kPlocRadius + k;
        if (idx >= 0 && idx < (int)m) {
            uint32_t c = clusters[idx];
            float4 l = blo[c], h = bhi[c];
            box[0][k] = l.x; box[1][k] = l.y; box[2][k] = l.z;
            box[3][k] = h.x; box[4][k] = h.y; box[5][k] = h.z;
        }
    }
    __syncthreads();
    const int i = base + (int)threadIdx.x;
    if (i >= (int)m) return;
    const int me = (int)threadIdx.x + kPlocRadius;
    const float ax = box[0][me], ay = box[1][me], az = box[2][me], bx = box[3][me], by = box[4][me], bz = box[5][me];
    float best = INFINITY;
    int bj = -1;
    const int j0 = max(0, i - kPlocRadius), j1 = min((int)m - 1, i + kPlocRadius);
    for (int j = j0; j <= j1; j++) {
        if (j == i) continue;
        const int s = j - base + kPlocRadius;
        float d = union_half_area(ax, ay, az, bx, by, bz, box[0][s], box[1][s], box[2][s], box[3][s], box[4][s], box[5][s]);
        if (bj < 0 || d < best) {   // j ascends: an equal area keeps the lower position
            best = d;
            bj = j;
        }
    }
    nn[i] = (uint32_t)bj;
}

// Per position: bit 0 = the cluster stays (it is not the higher partner of a
// mutual pair), bit 32 = it is the lower partner of a mutual pair (a merge).
__global__ void k_ploc_flags(const uint32_t* __restrict__ nn, uint32_t m, unsigned long long* __restrict__ flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t j = nn[i];
    bool mutual = nn[j] == i;
    unsigned long long stay = (mutual && j < i) ? 0ull : 1ull;
    unsigned long long merge = (mutual && i < j) ? 1ull : 0ull;
    flags[i] = stay | (merge << 32);
}

// Writes the next cluster array; a merging cluster becomes the new node
// next_id + (merges before it), with its union box and leaf count.
__global__ void k_ploc_compact(const uint32_t* __restrict__ clusters, const uint32_t* __restrict__ nn, uint32_t m,
                               const unsigned long long* __restrict__ flags, const unsigned long long* __restrict__ scan,
                               uint32_t n, uint32_t next_id, float4* __restrict__ blo, float4* __restrict__ bhi,
                               int2* __restrict__ child, uint32_t* __restrict__ parent, uint32_t* __restrict__ count,
                               uint32_t* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    unsigned long long f = flags[i];
    if (!(f & 1ull)) return;
    unsigned long long sc = scan[i];
    uint32_t pos = (uint32_t)sc;
    uint32_t c = clusters[i];
    if (f >> 32) {
        uint32_t cj = clusters[nn[i]];
        uint32_t id = next_id + (uint32_t)(sc >> 32);
        float4 al = blo[c], ah = bhi[c], bl = blo[cj], bh = bhi[cj];
        blo[id] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.f);
        bhi[id] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.f);
        child[id - n] = make_int2((int)c, (int)cj);
        parent[c] = id;
        parent[cj] = id;
        count[id] = count[c] + count[cj];
        out[pos] = id;
    } else {
        out[pos] = c;
    }
}

// Position of every node's first leaf in the depth-first leaf order (lower
// child first): the leaf counts of the lower siblings of the node's
// ancestors (and of the node itself) that it lies to the right of.
__global__ void k_ploc_offsets(uint32_t n, const int2* __restrict__ child, const uint32_t* __restrict__ parent,
                               const uint32_t* __restrict__ count, uint32_t* __restrict__ off) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t root = 2 * n - 2;
    if (x > root) return;
    uint32_t o = 0, y = x;
    while (y != root) {
        uint32_t p = parent[y];
        int2 c = child[p - n];
        if ((uint32_t)c.y == y) o += count[(uint32_t)c.x];
        y = p;
    }
    off[x] = o;
}

// The clustered tree in the radix tree's form: internal node x -> 2n-2-x
// (root 0), leaf x -> binary ref n-1+off[x], and the leaf order.
__global__ void k_ploc_to_radix(uint32_t n, const int2* __restrict__ pchild, const uint32_t* __restrict__ pparent,
                                const uint32_t* __restrict__ pcount, const uint32_t* __restrict__ off,
                                const uint32_t* __restrict__ sorted, uint32_t* __restrict__ order, int2* __restrict__ childs,
                                int2* __restrict__ range, uint32_t* __restrict__ parent) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t root = 2 * n - 2;
    if (x > root) return;
    auto ref = [&](uint32_t y) { return y < n ? (n - 1 + off[y]) : (root - y); };
    if (x < n) {
        order[off[x]] = sorted[x];
    } else {
        int2 c = pchild[x - n];
        childs[root - x] = make_int2((int)ref((uint32_t)c.x), (int)ref((uint32_t)c.y));
        range[root - x] = make_int2((int)off[x], (int)(off[x] + pcount[x] - 1));
    }
    if (x != root) parent[ref(x)] = root - pparent[x];
}

// ---------------------------------------------------------------------------
// Binned SAH on the GPU (bvh_builder 3): the host builder's splits, computed
// on the device, so the trees are the same. Ranges of more than kSahSmall
// primitives are split level by level by block-wide passes over chunks of at
// most kSahChunk positions: centroid bounds and bins (LDS, then global
// atomics on order-preserving integer encodings of the floats, so min and
// max are exact), the split choice (one wave per range), and a stable
// partition (left counts per chunk, a scan over each range's chunks, block
// prefix sums) into a second order buffer and back. Ranges of 2..kSahSmall
// primitives are then built whole by one wave each, with the range's order
// and its bins in LDS, children taken depth first (the larger one stacked,
// so the stack holds at most log2(kSahSmall) entries).
//
// The split choice runs the host's two sweeps over the 32 bins of an axis as
// inclusive scans across 32 lanes (sums of counts and min / max of bounds are
// exact in any order), then takes the smallest cost, ties to the lowest
// (axis, bin): the first strict minimum of the host's loop.
// ---------------------------------------------------------------------------
constexpr int kSahSmall = 1024;
constexpr int kSahChunk = 4096;
constexpr int kSahBinWords = 3 * kSahBins * 7;   // per axis and bin: count, lo xyz, hi xyz (encoded)
constexpr int kSahStack = 16;                     // > log2(kSahSmall)

struct SahSplit {
    int axis, bin, m, mid;   // mid: split at the middle of the range (no partition)
    float cl, scale;         // bin arithmetic of the split axis
    int pad0, pad1;
};

__device__ inline float hmin(float a, float b) { return b < a ? b : a; }   // std::min
__device__ inline float hmax(float a, float b) { return a < b ? b : a; }   // std::max

__global__ void k_sah_centroids(const float4* __restrict__ lo, const float4* __restrict__ hi, uint32_t n,
                                float4* __restrict__ cen, uint32_t* __restrict__ order) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 l = lo[i], h = hi[i];
    cen[i] = make_float4(0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z), 0.f);
    order[i] = i;
}

__device__ inline void sah_bins_init(uint32_t* w, int i) {
    const int f = i % 7;
    w[i] = f == 0 ? 0u : (f < 4 ? f2o(INFINITY) : f2o(-INFINITY));
}

// Centroid bounds (6 encoded words) and bins of every range of a level.
__global__ void k_sah_level_init(uint32_t nn, uint32_t* __restrict__ cb, uint32_t* __restrict__ bins) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn * 6) cb[i] = (i % 6) < 3 ? f2o(INFINITY) : f2o(-INFINITY);
    if (i < nn * kSahBinWords) sah_bins_init(bins, (int)i);
}

__global__ void __launch_bounds__(256) k_sah_bounds(const int4* __restrict__ chunks, const uint32_t* __restrict__ order,
                                                    const float4* __restrict__ cen, uint32_t* __restrict__ cb) {
    const int4 ch = chunks[blockIdx.x];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = ch.y + (int)threadIdx.x; i < ch.z; i += 256) {
        float4 c = cen[order[i]];
        mn[0] = fminf(mn[0], c.x); mn[1] = fminf(mn[1], c.y); mn[2] = fminf(mn[2], c.z);
        mx[0] = fmaxf(mx[0], c.x); mx[1] = fmaxf(mx[1], c.y); mx[2] = fmaxf(mx[2], c.z);
    }
    for (int k = 0; k < 3; k++) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], off));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], off));
        }
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; k++) {
            atomicMin(&cb[6 * ch.x + k], f2o(mn[k]));
            atomicMax(&cb[6 * ch.x + 3 + k], f2o(mx[k]));
        }
}

// Adds one primitive to the encoded bins `w` of its range.
__device__ inline void sah_bin_prim(uint32_t* w, const float cl[3], const float scale[3], const bool live[3], float4 c,
                                    float4 l, float4 h) {
    const float cc[3] = {c.x, c.y, c.z};
    for (int k = 0; k < 3; k++) {
        if (!live[k]) continue;
        uint32_t* b = w + (k * kSahBins + sah_bin(cc[k], cl[k], scale[k])) * 7;
        atomicAdd(&b[0], 1u);
        atomicMin(&b[1], f2o(l.x)); atomicMin(&b[2], f2o(l.y)); atomicMin(&b[3], f2o(l.z));
        atomicMax(&b[4], f2o(h.x)); atomicMax(&b[5], f2o(h.y)); atomicMax(&b[6], f2o(h.z));
    }
}

__device__ inline void sah_axes(const float cl[3], const float ch[3], float scale[3], bool live[3]) {
    for (int k = 0; k < 3; k++) {
        const float ext = ch[k] - cl[k];
        live[k] = ext > 0.0f;
        scale[k] = live[k] ? float(kSahBins) / ext : 0.0f;
    }
}

__global__ void __launch_bounds__(256) k_sah_bin(const int4* __restrict__ chunks, const uint32_t* __restrict__ order,
                                                 const float4* __restrict__ cen, const float4* __restrict__ lo,
                                                 const float4* __restrict__ hi, const uint32_t* __restrict__ cb,
                                                 uint32_t* __restrict__ bins) {
    __shared__ uint32_t w[kSahBinWords];
    const int4 ch = chunks[blockIdx.x];
    for (int i = threadIdx.x; i < kSahBinWords; i += 256) sah_bins_init(w, i);
    float cl[3], cu[3], scale[3];
    bool live[3];
    for (int k = 0; k < 3; k++) {
        cl[k] = o2f(cb[6 * ch.x + k]);
        cu[k] = o2f(cb[6 * ch.x + 3 + k]);
    }
    sah_axes(cl, cu, scale, live);
    __syncthreads();
    for (int i = ch.y + (int)threadIdx.x; i < ch.z; i += 256) {
        uint32_t p = order[i];
        sah_bin_prim(w, cl, scale, live, cen[p], lo[p], hi[p]);
    }
    __syncthreads();
    uint32_t* g = bins + (size_t)ch.x * kSahBinWords;
    for (int i = threadIdx.x; i < kSahBinWords; i += 256) {
        const uint32_t v = w[i];
        const int f = i % 7;
        if (f == 0) {
            if (v) atomicAdd(&g[i], v);
        } else if (f < 4) {
            if (v != f2o(INFINITY)) atomicMin(&g[i], v);
        } else if (v != f2o(-INFINITY)) {
            atomicMax(&g[i], v);
        }
    }
}

// The split of one range from its encoded bins `w` (LDS) and its centroid
// bounds, by the 64 lanes of one wave; every lane returns the same choice.
// axis < 0: no split with a finite cost (coincident centroids).
struct SahChoice {
    int axis, bin, nleft;
};
__device__ SahChoice sah_choose_wave(const uint32_t* w, const float cl[3], const float cu[3]) {
    const int lane = (int)__lane_id();
    const int q = lane & (kSahBins - 1);
    const bool act = lane < kSahBins;
    float best = INFINITY;
    int bidx = -1;
    for (int k = 0; k < 3; k++) {
        const float ext = cu[k] - cl[k];
        if (!(ext > 0.0f)) continue;
        const uint32_t* b = w + (k * kSahBins + q) * 7;
        const int cnt = act ? (int)b[0] : 0;
        const float lx = act ? o2f(b[1]) : INFINITY, ly = act ? o2f(b[2]) : INFINITY, lz = act ? o2f(b[3]) : INFINITY;
        const float hx = act ? o2f(b[4]) : -INFINITY, hy = act ? o2f(b[5]) : -INFINITY, hz = act ? o2f(b[6]) : -INFINITY;
        // right side: bins q..31 (suffix), left side: bins 0..q (prefix)
        int rc = cnt, lc = cnt;
        float rlx = lx, rly = ly, rlz = lz, rhx = hx, rhy = hy, rhz = hz;
        float llx = lx, lly = ly, llz = lz, lhx = hx, lhy = hy, lhz = hz;
        for (int off = 1; off < kSahBins; off <<= 1) {
            const int orc = __shfl_down(rc, off);
            const float olx = __shfl_down(rlx, off), oly = __shfl_down(rly, off), olz = __shfl_down(rlz, off);
            const float ohx = __shfl_down(rhx, off), ohy = __shfl_down(rhy, off), ohz = __shfl_down(rhz, off);
            if (q + off < kSahBins) {
                rc += orc;
                rlx = hmin(rlx, olx); rly = hmin(rly, oly); rlz = hmin(rlz, olz);
                rhx = hmax(rhx, ohx); rhy = hmax(rhy, ohy); rhz = hmax(rhz, ohz);
            }
            const int plc = __shfl_up(lc, off);
            const float plx = __shfl_up(llx, off), ply = __shfl_up(lly, off), plz = __shfl_up(llz, off);
            const float phx = __shfl_up(lhx, off), phy = __shfl_up(lhy, off), phz = __shfl_up(lhz, off);
            if (q >= off) {
                lc += plc;
                llx = hmin(llx, plx); lly = hmin(lly, ply); llz = hmin(llz, plz);
                lhx = hmax(lhx, phx); lhy = hmax(lhy, phy); lhz = hmax(lhz, phz);
            }
        }
        const float rarea = rc ? sah_area(rlx, rly, rlz, rhx, rhy, rhz) : 0.0f;
        const int rc_next = __shfl_down(rc, 1);
        const float ra_next = __shfl_down(rarea, 1);
        float cost = INFINITY;
        if (act && q < kSahBins - 1 && lc != 0 && rc_next != 0)
            cost = sah_cost(lc, sah_area(llx, lly, llz, lhx, lhy, lhz), rc_next, ra_next);
        if (!(cost < INFINITY)) cost = INFINITY;   // NaN or infinite: never chosen
        int bq = cost < INFINITY ? q : 1 << 30;
        for (int off = 32; off > 0; off >>= 1) {
            const float oc = __shfl_xor(cost, off);
            const int oq = __shfl_xor(bq, off);
            if (oc < cost || (oc == cost && oq < bq)) {
                cost = oc;
                bq = oq;
            }
        }
        if (bq < (1 << 30) && cost < best) {   // strict: an earlier axis keeps a tie
            best = cost;
            bidx = k * kSahBins + bq;
        }
    }
    SahChoice r{-1, -1, 0};
    if (bidx >= 0) {
        r.axis = bidx / kSahBins;
        r.bin = bidx % kSahBins;
        int v = (act && q <= r.bin) ? (int)w[(r.axis * kSahBins + q) * 7] : 0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        r.nleft = v;
    }
    return r;
}

__global__ void __launch_bounds__(64) k_sah_choose(const int4* __restrict__ nodes, const uint32_t* __restrict__ cb,
                                                   const uint32_t* __restrict__ bins, SahSplit* __restrict__ splits) {
    __shared__ uint32_t w[kSahBinWords];
    const int j = blockIdx.x;
    for (int i = threadIdx.x; i < kSahBinWords; i += 64) w[i] = bins[(size_t)j * kSahBinWords + i];
    __syncthreads();
    float cl[3], cu[3];
    for (int k = 0; k < 3; k++) {
        cl[k] = o2f(cb[6 * j + k]);
        cu[k] = o2f(cb[6 * j + 3 + k]);
    }
    const SahChoice c = sah_choose_wave(w, cl, cu);
    if (threadIdx.x == 0) {
        const int4 nd = nodes[j];
        SahSplit s{};
        s.axis = c.axis;
        s.bin = c.bin;
        s.m = nd.x + c.nleft;
        s.mid = 0;
        if (c.axis < 0 || !(s.m > nd.x && s.m < nd.y)) {
            s.mid = 1;
            s.m = nd.x + (nd.y - nd.x) / 2;
        } else {
            s.cl = cl[c.axis];
            s.scale = float(kSahBins) / (cu[c.axis] - cl[c.axis]);
        }
        splits[j] = s;
    }
}

__device__ inline bool sah_goes_left(const SahSplit& s, int i, const float4* cen, uint32_t p) {
    if (s.mid) return i < s.m;
    const float4 c = cen[p];
    const float v = s.axis == 0 ? c.x : (s.axis == 1 ? c.y : c.z);
    return sah_bin(v, s.cl, s.scale) <= s.bin;
}

__global__ void __launch_bounds__(256) k_sah_count(const int4* __restrict__ chunks, const SahSplit* __restrict__ splits,
                                                   const uint32_t* __restrict__ order, const float4* __restrict__ cen,
                                                   uint32_t* __restrict__ chunk_left) {
    __shared__ uint32_t part[4];
    const int4 ch = chunks[blockIdx.x];
    const SahSplit s = splits[ch.x];
    uint32_t cnt = 0;
    for (int i = ch.y + (int)threadIdx.x; i < ch.z; i += 256) cnt += sah_goes_left(s, i, cen, order[i]) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) chunk_left[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Per range: the left count before each of its chunks.
__global__ void k_sah_chunkscan(uint32_t nn, const int2* __restrict__ node_chunks, const uint32_t* __restrict__ chunk_left,
                                uint32_t* __restrict__ chunk_base) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nn) return;
    const int2 nc = node_chunks[j];
    uint32_t run = 0;
    for (int c = nc.x; c < nc.x + nc.y; c++) {
        chunk_base[c] = run;
        run += chunk_left[c];
    }
}

// Stable partition of each chunk's positions into `out` (left side from the
// range's start, right side from its split position).
__global__ void __launch_bounds__(256) k_sah_scatter(const int4* __restrict__ chunks, const int4* __restrict__ nodes,
                                                     const SahSplit* __restrict__ splits, const uint32_t* __restrict__ order,
                                                     const float4* __restrict__ cen, const uint32_t* __restrict__ chunk_base,
                                                     uint32_t* __restrict__ out) {
    __shared__ uint32_t wsum[4];
    const int4 ch = chunks[blockIdx.x];
    const int b = nodes[ch.x].x;
    const SahSplit s = splits[ch.x];
    const int lane = (int)__lane_id(), wv = (int)threadIdx.x >> 6;
    uint32_t run = chunk_base[blockIdx.x];
    for (int base = ch.y; base < ch.z; base += 256) {
        const int i = base + (int)threadIdx.x;
        const bool in = i < ch.z;
        const uint32_t p = in ? order[i] : 0u;
        const bool left = in && sah_goes_left(s, i, cen, p);
        const uint64_t bal = __ballot(left);
        const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wv] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t before = run;
        for (int k = 0; k < wv; k++) before += wsum[k];
        const uint32_t total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
        if (in) {
            const uint32_t lr = before + below;   // left flags in [b, i)
            const int pos = left ? b + (int)lr : s.m + (i - b - (int)lr);
            out[pos] = p;
        }
        run += total;
    }
}

__global__ void __launch_bounds__(256) k_sah_copyback(const int4* __restrict__ chunks, const uint32_t* __restrict__ src,
                                                      uint32_t* __restrict__ dst) {
    const int4 ch = chunks[blockIdx.x];
    for (int i = ch.y + (int)threadIdx.x; i < ch.z; i += 256) dst[i] = src[i];
}

// Children of the binary node `id` over [b, e) split at m (the host
// builder's numbering: left child id + 1, right child id + (m - b), a
// one-primitive side is the leaf ref n - 1 + position).
__device__ inline void sah_emit_node(int b, int e, int id, int m, int n, int2* childs, int2* range, uint32_t* parent) {
    const int lid = id + 1, rid = id + (m - b);
    const int l = m - b == 1 ? n - 1 + b : lid;
    const int r = e - m == 1 ? n - 1 + m : rid;
    childs[id] = make_int2(l, r);
    range[id] = make_int2(b, e - 1);
    parent[l] = (uint32_t)id;
    parent[r] = (uint32_t)id;
}

// Ranges are (first position, end position, binary node id, mesh); a
// mesh's primitives hold positions P..P+n-1 (meshes[mesh] = (P, n)), and its
// tree is written in its own numbering to childs + P, range + P and
// parent + 2P.
__device__ inline void sah_emit_mesh(int4 nd, int m, const int2* meshes, int2* childs, int2* range, uint32_t* parent) {
    const int2 mp = meshes[nd.w];
    sah_emit_node(nd.x - mp.x, nd.y - mp.x, nd.z, m - mp.x, mp.y, childs + mp.x, range + mp.x, parent + 2 * (size_t)mp.x);
}

__global__ void k_sah_emit(uint32_t nn, const int4* __restrict__ nodes, const SahSplit* __restrict__ splits,
                           const int2* __restrict__ meshes, int2* __restrict__ childs, int2* __restrict__ range,
                           uint32_t* __restrict__ parent, int4* __restrict__ next, uint32_t* __restrict__ next_count,
                           int4* __restrict__ small, uint32_t* __restrict__ small_count) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nn) return;
    const int4 nd = nodes[j];
    const int m = splits[j].m;
    sah_emit_mesh(nd, m, meshes, childs, range, parent);
    const int4 kids[2] = {make_int4(nd.x, m, nd.z + 1, nd.w), make_int4(m, nd.y, nd.z + (m - nd.x), nd.w)};
    for (int c = 0; c < 2; c++) {
        const int sz = kids[c].y - kids[c].x;
        if (sz > kSahSmall) next[atomicAdd(next_count, 1u)] = kids[c];
        else if (sz >= 2) small[atomicAdd(small_count, 1u)] = kids[c];
    }
}

// One wave builds the whole subtree over one range of 2..kSahSmall
// primitives. The range's centroids and boxes are staged in LDS once (slot j
// = the j-th primitive of the range), and the order is kept as slots.
__global__ void __launch_bounds__(64) k_sah_small(const int4* __restrict__ ranges, const int2* __restrict__ meshes,
                                                  const float4* __restrict__ cen,
                                                  const float4* __restrict__ lo, const float4* __restrict__ hi,
                                                  uint32_t* __restrict__ order, int2* __restrict__ childs,
                                                  int2* __restrict__ range, uint32_t* __restrict__ parent) {
    __shared__ uint16_t s_ord[kSahSmall], s_tmp[kSahSmall];
    __shared__ uint32_t s_gid[kSahSmall];
    __shared__ float s_c[3][kSahSmall], s_lo[3][kSahSmall], s_hi[3][kSahSmall];
    __shared__ uint32_t w[kSahBinWords];
    __shared__ int4 stack[kSahStack];
    const int lane = (int)threadIdx.x;
    const int4 root = ranges[blockIdx.x];
    const int b0 = root.x, S = root.y - root.x;
    for (int j = lane; j < S; j += 64) {
        const uint32_t p = order[b0 + j];
        const float4 c = cen[p], l = lo[p], h = hi[p];
        s_gid[j] = p;
        s_ord[j] = (uint16_t)j;
        s_c[0][j] = c.x; s_c[1][j] = c.y; s_c[2][j] = c.z;
        s_lo[0][j] = l.x; s_lo[1][j] = l.y; s_lo[2][j] = l.z;
        s_hi[0][j] = h.x; s_hi[1][j] = h.y; s_hi[2][j] = h.z;
    }
    __syncthreads();
    int cb = root.x, ce = root.y, cid = root.z, sp = 0;
    for (;;) {
        // cb, ce, cid, sp are uniform across the wave
        if (ce - cb >= 2) {
            float cl[3] = {INFINITY, INFINITY, INFINITY}, cu[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int j = cb - b0 + lane; j < ce - b0; j += 64) {
                const int t = s_ord[j];
                for (int k = 0; k < 3; k++) {
                    cl[k] = fminf(cl[k], s_c[k][t]);
                    cu[k] = fmaxf(cu[k], s_c[k][t]);
                }
            }
            for (int k = 0; k < 3; k++)
                for (int off = 32; off > 0; off >>= 1) {
                    cl[k] = fminf(cl[k], __shfl_xor(cl[k], off));
                    cu[k] = fmaxf(cu[k], __shfl_xor(cu[k], off));
                }
            float scale[3];
            bool live[3];
            sah_axes(cl, cu, scale, live);
            for (int i = lane; i < kSahBinWords; i += 64) sah_bins_init(w, i);
            __syncthreads();
            for (int j = cb - b0 + lane; j < ce - b0; j += 64) {
                const int t = s_ord[j];
                sah_bin_prim(w, cl, scale, live, make_float4(s_c[0][t], s_c[1][t], s_c[2][t], 0.f),
                             make_float4(s_lo[0][t], s_lo[1][t], s_lo[2][t], 0.f),
                             make_float4(s_hi[0][t], s_hi[1][t], s_hi[2][t], 0.f));
            }
            __syncthreads();
            const SahChoice c = sah_choose_wave(w, cl, cu);
            int m = cb + c.nleft;
            const bool mid = c.axis < 0 || !(m > cb && m < ce);
            if (mid) {
                m = cb + (ce - cb) / 2;
            } else {
                const float scl = float(kSahBins) / (cu[c.axis] - cl[c.axis]);
                uint32_t run = 0;
                for (int base = cb; base < ce; base += 64) {
                    const int i = base + lane;
                    const bool in = i < ce;
                    const int t = in ? s_ord[i - b0] : 0;
                    const bool left = in && sah_bin(s_c[c.axis][t], cl[c.axis], scl) <= c.bin;
                    const uint64_t bal = __ballot(left);
                    const uint32_t lr = run + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                    if (in) s_tmp[(left ? cb + (int)lr : m + (i - cb - (int)lr)) - b0] = (uint16_t)t;
                    run += (uint32_t)__popcll(bal);
                }
                __syncthreads();
                for (int j = cb - b0 + lane; j < ce - b0; j += 64) s_ord[j] = s_tmp[j];
                __syncthreads();
            }
            if (lane == 0) sah_emit_mesh(make_int4(cb, ce, cid, root.w), m, meshes, childs, range, parent);
            const int ls = m - cb, rs = ce - m;
            const int4 L = make_int4(cb, m, cid + 1, 0), R = make_int4(m, ce, cid + ls, 0);
            if (ls >= 2 && rs >= 2) {
                const int4 big = ls >= rs ? L : R, sml = ls >= rs ? R : L;
                if (lane == 0) stack[sp] = big;
                sp++;
                cb = sml.x; ce = sml.y; cid = sml.z;
                __syncthreads();
                continue;
            }
            if (ls >= 2) { cb = L.x; ce = L.y; cid = L.z; continue; }
            if (rs >= 2) { cb = R.x; ce = R.y; cid = R.z; continue; }
        }
        if (sp == 0) break;
        sp--;
        const int4 t = stack[sp];
        cb = t.x; ce = t.y; cid = t.z;
    }
    __syncthreads();
    for (int j = lane; j < S; j += 64) order[b0 + j] = s_gid[s_ord[j]];
}

inline uint32_t blocks(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

// Builds a W-wide BVH over n primitive AABBs (device lo/hi float4 arrays).
// Nodes are written to `nodes` starting at index 0 of that array; child
// references are offset by node_base (internal) and leaf_base (leaves).
// On return *order_out (device, n entries) holds the primitive index of every
// leaf slot and root_box the bounds of all primitives.
// ---------------------------------------------------------------------------
// Binned SAH build on the host: the quality a driver's PREFER_FAST_TRACE
// build aims for (mesh.cpp:109), at the cost of a host pass per mesh. It
// produces the same binary-tree form as the radix tree (internal nodes
// 0..n-2, leaf j = binary ref n-1+j over the primitive order `order`,
// contiguous ranges, parents), so the GPU refit and the wide collapse are
// shared. Every split is chosen among 32 centroid bins on each axis by the
// surface-area cost (a split costs its children's area-weighted counts),
// and the range is partitioned stably, so every range stays in ascending
// primitive order; ranges whose centroids coincide are split in the middle.
// The subtree over m primitives owns the m - 1 internal ids after its root,
// so subtrees are built in parallel without shared state.
// ---------------------------------------------------------------------------
namespace {
struct SahPrim {
    float lo[3], hi[3], c[3];
};

struct SahTree {
    const std::vector<SahPrim>* prims;
    std::vector<uint32_t> order;
    std::vector<int2> childs, range;
    std::vector<uint32_t> parent;
    int n = 0;

    static float area(const float lo[3], const float hi[3]) { return sah_area(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]); }

    // split [b, e) and return the first index of the right side
    int split(int b, int e) {
        const std::vector<SahPrim>& P = *prims;
        float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = b; i < e; i++)
            for (int k = 0; k < 3; k++) {
                cl[k] = std::min(cl[k], P[order[i]].c[k]);
                ch[k] = std::max(ch[k], P[order[i]].c[k]);
            }
        constexpr int NB = kSahBins;
        float best = INFINITY;
        int best_axis = -1, best_bin = -1;
        for (int k = 0; k < 3; k++) {
            const float ext = ch[k] - cl[k];
            if (!(ext > 0.0f)) continue;
            const float scale = NB / ext;
            int cnt[NB] = {0};
            float blo[NB][3], bhi[NB][3];
            for (int q = 0; q < NB; q++)
                for (int a = 0; a < 3; a++) { blo[q][a] = INFINITY; bhi[q][a] = -INFINITY; }
            for (int i = b; i < e; i++) {
                const SahPrim& p = P[order[i]];
                const int q = sah_bin(p.c[k], cl[k], scale);
                cnt[q]++;
                for (int a = 0; a < 3; a++) {
                    blo[q][a] = std::min(blo[q][a], p.lo[a]);
                    bhi[q][a] = std::max(bhi[q][a], p.hi[a]);
                }
            }
            float rarea[NB];
            int rcnt[NB];
            float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
            int rc = 0;
            for (int q = NB - 1; q > 0; q--) {
                rc += cnt[q];
                for (int a = 0; a < 3; a++) { rl[a] = std::min(rl[a], blo[q][a]); rh[a] = std::max(rh[a], bhi[q][a]); }
                rcnt[q] = rc;
                rarea[q] = rc ? area(rl, rh) : 0.0f;
            }
            float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
            int lc = 0;
            for (int q = 0; q < NB - 1; q++) {
                lc += cnt[q];
                for (int a = 0; a < 3; a++) { ll[a] = std::min(ll[a], blo[q][a]); lh[a] = std::max(lh[a], bhi[q][a]); }
                if (lc == 0 || rcnt[q + 1] == 0) continue;
                const float cost = sah_cost(lc, area(ll, lh), rcnt[q + 1], rarea[q + 1]);
                if (cost < best) { best = cost; best_axis = k; best_bin = q; }
            }
        }
        if (best_axis >= 0) {
            const float ext = ch[best_axis] - cl[best_axis], scale = NB / ext;
            // stable: every range stays in ascending primitive order
            auto mid = std::stable_partition(order.begin() + b, order.begin() + e, [&](uint32_t i) {
                return sah_bin(P[i].c[best_axis], cl[best_axis], scale) <= best_bin;
            });
            const int m = int(mid - order.begin());
            if (m > b && m < e) return m;
        }
        // coincident centroids (or no finite cost): the middle of the range,
        // which is in ascending primitive order
        return b + (e - b) / 2;
    }

    int ref_of(int b, int e, int id) const { return e - b == 1 ? n - 1 + b : id; }

    void build(int b, int e, int id, int depth) {
        const int m = split(b, e);
        const int lid = id + 1, rid = id + (m - b);
        const int l = ref_of(b, m, lid), r = ref_of(m, e, rid);
        childs[id] = make_int2(l, r);
        range[id] = make_int2(b, e - 1);
        parent[l] = uint32_t(id);
        parent[r] = uint32_t(id);
        const bool go_left = m - b > 1, go_right = e - m > 1;
        // the top five levels hand one side to a thread of its own (up to 32
        // threads); the node ids do not depend on the schedule
        if (go_left && go_right && depth < 5 && e - b > 4096) {
            std::thread t([&, b, m, lid, depth] { build(b, m, lid, depth + 1); });
            build(m, e, rid, depth + 1);
            t.join();
            return;
        }
        if (go_left) build(b, m, lid, depth + 1);
        if (go_right) build(m, e, rid, depth + 1);
    }
};
}  // namespace

// Host tree over n primitive boxes (lo/hi, 3 floats each) -> order, childs,
// range, parent in the radix tree's form.
static void sah_tree(uint32_t n, const float* lo, const float* hi, SahTree& t, std::vector<SahPrim>& prims) {
    prims.resize(n);
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            prims[i].lo[k] = lo[3 * i + k];
            prims[i].hi[k] = hi[3 * i + k];
            prims[i].c[k] = 0.5f * (lo[3 * i + k] + hi[3 * i + k]);
        }
    t.prims = &prims;
    t.n = int(n);
    t.order.resize(n);
    for (uint32_t i = 0; i < n; i++) t.order[i] = i;
    t.childs.assign(std::max<uint32_t>(n, 1), make_int2(0, 0));
    t.range.assign(std::max<uint32_t>(n, 1), make_int2(0, 0));
    t.parent.assign(2 * size_t(std::max<uint32_t>(n, 1)), 0u);
    if (n > 1) t.build(0, int(n), 0, 0);
}

static hipError_t sah_forest_gpu(hipStream_t st, const std::vector<int2>& meshes, uint32_t total, const float4* d_lo,
                                 const float4* d_hi, uint32_t* d_order, int2* childs, int2* range, uint32_t* parent,
                                 BuildScratch& sc);

// Bounds (refit) and the W-wide collapse of one binary tree in the radix
// tree's form over n primitives: `order` maps leaf positions to indices into
// d_lo / d_hi; childs / range / flags hold n - 1 entries, parent / blo / bhi
// 2n - 1. Nodes go to d_nodes (global index node_base + i), leaves reference
// leaf_base + position (or the primitive when leaf_is_prim).
template <int W>
static hipError_t finish_wide(hipStream_t st, uint32_t n, const float4* d_lo, const float4* d_hi, const uint32_t* d_order,
                              const int2* childs, const int2* range, const uint32_t* parent, uint32_t* flags, float4* blo,
                              float4* bhi, BuildScratch& sc, int leaf_size, int leaf_is_prim, WideNode<W>* d_nodes,
                              uint32_t node_base, uint32_t leaf_base, uint32_t* num_nodes, float root_box[6]) {
    const uint32_t B = 256;
    k_refit<<<blocks(n, B), B, 0, st>>>(d_lo, d_hi, d_order, (int)n, childs, parent, flags, blo, bhi);
    int2 job0 = make_int2(0, 0);   // binary ref of the root (leaf 0 when n == 1: ref n-1+0 = 0)
    uint32_t one = 1;
    HH_TRY(hipMemcpyAsync(sc.jobs_a, &job0, sizeof(int2), hipMemcpyHostToDevice, st));
    HH_TRY(hipMemcpyAsync(sc.counters, &one, 4, hipMemcpyHostToDevice, st));       // jobs in A
    HH_TRY(hipMemcpyAsync(sc.counters + 2, &one, 4, hipMemcpyHostToDevice, st));   // node counter
    uint32_t njobs = 1;
    int2* qa = sc.jobs_a;
    int2* qb = sc.jobs_b;
    uint32_t* ca = sc.counters;
    uint32_t* cb = sc.counters + 1;
    while (njobs > 0) {
        k_fill_u32<<<1, 1, 0, st>>>(cb, 0u, 1);
        k_collapse<W><<<blocks(njobs, 64), 64, 0, st>>>(qa, ca, qb, cb, sc.counters + 2, (int)n, leaf_size, childs, range, blo,
                                                         bhi, d_order, d_nodes, node_base, leaf_base, leaf_is_prim);
        HH_TRY(hipMemcpyAsync(&njobs, cb, 4, hipMemcpyDeviceToHost, st));
        HH_TRY(hipStreamSynchronize(st));
        std::swap(qa, qb);
        std::swap(ca, cb);
    }
    HH_TRY(hipMemcpyAsync(num_nodes, sc.counters + 2, 4, hipMemcpyDeviceToHost, st));
    float4 rb[2];
    HH_TRY(hipMemcpyAsync(&rb[0], &blo[0], 16, hipMemcpyDeviceToHost, st));
    HH_TRY(hipMemcpyAsync(&rb[1], &bhi[0], 16, hipMemcpyDeviceToHost, st));
    HH_TRY(hipStreamSynchronize(st));
    root_box[0] = rb[0].x; root_box[1] = rb[0].y; root_box[2] = rb[0].z;
    root_box[3] = rb[1].x; root_box[4] = rb[1].y; root_box[5] = rb[1].z;
    return hipSuccess;
}

template <int W>
static hipError_t build_wide(hipStream_t st, uint32_t n, const float4* d_lo, const float4* d_hi, const uint32_t* d_cbounds,
                             WideNode<W>* d_nodes, uint32_t node_base, uint32_t leaf_base, int leaf_size, int leaf_is_prim,
                             uint32_t* d_order, BuildScratch& sc, uint32_t* num_nodes, float root_box[6],
                             const float* h_lo = nullptr, const float* h_hi = nullptr, bool ploc = false,
                             bool sah_gpu = false) {
    HH_TRY(sc.ensure(n));
    const uint32_t B = 256;
    if (sah_gpu && n > 1) {
        HH_TRY(sc.ensure_sah(n));
        const std::vector<int2> one_mesh{make_int2(0, (int)n)};
        HH_TRY(sah_forest_gpu(st, one_mesh, n, d_lo, d_hi, d_order, sc.childs, sc.range, sc.parent, sc));
        k_fill_u32<<<blocks(n, B), B, 0, st>>>(sc.flags, 0u, n);
    } else if (sah_gpu) {
        k_fill_u32<<<1, 1, 0, st>>>(d_order, 0u, 1);
    } else if (ploc && n > 1) {
        HH_TRY(sc.ensure_ploc(n));
        k_morton<<<blocks(n, B), B, 0, st>>>(d_lo, d_hi, n, d_cbounds, sc.keys_in, sc.vals_in);
        size_t tb = sc.sort_bytes;
        HH_TRY(hipcub::DeviceRadixSort::SortPairs(sc.sort_tmp, tb, sc.keys_in, sc.keys_out, sc.vals_in, sc.p_sorted, (int)n, 0,
                                                  30, st));
        // PLOC boxes live in blo/bhi (ids 0..2n-2); the refit below rewrites
        // them in the radix tree's numbering.
        k_ploc_init<<<blocks(n, B), B, 0, st>>>(d_lo, d_hi, sc.p_sorted, n, sc.blo, sc.bhi, sc.p_cl[0], sc.p_count);
        uint32_t m = n, next_id = n, cur = 0;
        unsigned long long* tail = sc.p_tail;
        while (m > 1) {
            k_ploc_nn<<<blocks(m, kPlocBlock), kPlocBlock, 0, st>>>(sc.p_cl[cur], m, sc.blo, sc.bhi, sc.p_nn);
            k_ploc_flags<<<blocks(m, B), B, 0, st>>>(sc.p_nn, m, sc.p_flags);
            size_t sb = sc.scan_bytes;
            HH_TRY(hipcub::DeviceScan::ExclusiveSum(sc.scan_tmp, sb, sc.p_flags, sc.p_scan, (int)m, st));
            k_ploc_compact<<<blocks(m, B), B, 0, st>>>(sc.p_cl[cur], sc.p_nn, m, sc.p_flags, sc.p_scan, n, next_id, sc.blo,
                                                       sc.bhi, sc.p_child, sc.p_parent, sc.p_count, sc.p_cl[cur ^ 1]);
            HH_TRY(hipMemcpyAsync(&tail[0], sc.p_flags + (m - 1), 8, hipMemcpyDeviceToHost, st));
            HH_TRY(hipMemcpyAsync(&tail[1], sc.p_scan + (m - 1), 8, hipMemcpyDeviceToHost, st));
            HH_TRY(hipStreamSynchronize(st));
            const unsigned long long tot = tail[0] + tail[1];
            const uint32_t stay = (uint32_t)tot, merges = (uint32_t)(tot >> 32);
            if (merges == 0 || stay + merges != m) return hipErrorUnknown;   // every round merges a pair
            m = stay;
            next_id += merges;
            cur ^= 1;
        }
        if (next_id != 2 * n - 1) return hipErrorUnknown;
        k_ploc_offsets<<<blocks(2 * n - 1, B), B, 0, st>>>(n, sc.p_child, sc.p_parent, sc.p_count, sc.p_off);
        k_ploc_to_radix<<<blocks(2 * n - 1, B), B, 0, st>>>(n, sc.p_child, sc.p_parent, sc.p_count, sc.p_off, sc.p_sorted, d_order,
                                                           sc.childs, sc.range, sc.parent);
        k_fill_u32<<<blocks(n, B), B, 0, st>>>(sc.flags, 0u, n);
    } else if (ploc) {
        k_fill_u32<<<1, 1, 0, st>>>(d_order, 0u, 1);
    } else if (h_lo) {
        // binned SAH on the host (the boxes are the host copies of d_lo / d_hi)
        SahTree t;
        std::vector<SahPrim> prims;
        sah_tree(n, h_lo, h_hi, t, prims);
        HH_TRY(hipMemcpyAsync(d_order, t.order.data(), size_t(n) * 4, hipMemcpyHostToDevice, st));
        if (n > 1) {
            HH_TRY(hipMemcpyAsync(sc.childs, t.childs.data(), size_t(n - 1) * sizeof(int2), hipMemcpyHostToDevice, st));
            HH_TRY(hipMemcpyAsync(sc.range, t.range.data(), size_t(n - 1) * sizeof(int2), hipMemcpyHostToDevice, st));
            HH_TRY(hipMemcpyAsync(sc.parent, t.parent.data(), (2 * size_t(n) - 1) * 4, hipMemcpyHostToDevice, st));
            k_fill_u32<<<blocks(n, B), B, 0, st>>>(sc.flags, 0u, n);
        }
        HH_TRY(hipStreamSynchronize(st));   // the host vectors go out of scope
    } else {
        k_morton<<<blocks(n, B), B, 0, st>>>(d_lo, d_hi, n, d_cbounds, sc.keys_in, sc.vals_in);
        size_t tb = sc.sort_bytes;
        HH_TRY(hipcub::DeviceRadixSort::SortPairs(sc.sort_tmp, tb, sc.keys_in, sc.keys_out, sc.vals_in, d_order, (int)n, 0, 30, st));
        if (n > 1) {
            k_radix_tree<<<blocks(n - 1, B), B, 0, st>>>(sc.keys_out, (int)n, sc.childs, sc.parent, sc.range);
            k_fill_u32<<<blocks(n, B), B, 0, st>>>(sc.flags, 0u, n);
        }
    }
    return finish_wide<W>(st, n, d_lo, d_hi, d_order, sc.childs, sc.range, sc.parent, sc.flags, sc.blo, sc.bhi, sc, leaf_size,
                          leaf_is_prim, d_nodes, node_base, leaf_base, num_nodes, root_box);
}

hipError_t BuildScratch::ensure_sah(uint32_t n) {
    if (n <= s_cap) return hipSuccess;
    release_sah();
    s_cap = std::max<uint32_t>(n, 1024);
    const size_t N = s_cap;
    s_nmax = s_cap / kSahSmall + 2;                      // disjoint ranges of > kSahSmall primitives
    s_cmax = s_cap / kSahChunk + 2 * s_nmax + 2;         // chunks of <= kSahChunk positions per range
    HH_TRY(hipMalloc(&s_cen, N * sizeof(float4)));
    HH_TRY(hipMalloc(&s_ord2, N * 4));
    HH_TRY(hipMalloc(&s_cb, size_t(s_nmax) * 6 * 4));
    HH_TRY(hipMalloc(&s_bins, size_t(s_nmax) * kSahBinWords * 4));
    HH_TRY(hipMalloc(&s_splits, size_t(s_nmax) * sizeof(SahSplit)));
    HH_TRY(hipMalloc(&s_nodes, size_t(s_nmax) * sizeof(int4)));
    HH_TRY(hipMalloc(&s_next, size_t(s_nmax) * sizeof(int4)));
    HH_TRY(hipMalloc(&s_small, (N / 2 + 1) * sizeof(int4)));   // disjoint ranges of >= 2 primitives
    HH_TRY(hipMalloc(&s_chunks, size_t(s_cmax) * sizeof(int4)));
    HH_TRY(hipMalloc(&s_node_chunks, size_t(s_nmax) * sizeof(int2)));
    HH_TRY(hipMalloc(&s_chunk_left, size_t(s_cmax) * 4));
    HH_TRY(hipMalloc(&s_chunk_base, size_t(s_cmax) * 4));
    HH_TRY(hipMalloc(&s_ctr, 16));
    HH_TRY(hipHostMalloc(&s_tail, 16));
    return hipSuccess;
}

hipError_t BuildScratch::ensure_sah_meshes(uint32_t nm) {
    if (nm <= s_mesh_cap) return hipSuccess;
    if (s_meshes) (void)hipFree(s_meshes);
    s_meshes = nullptr;
    s_mesh_cap = std::max<uint32_t>(nm, 64);
    HH_TRY(hipMalloc(&s_meshes, size_t(s_mesh_cap) * sizeof(int2)));
    return hipSuccess;
}

void BuildScratch::release_sah() {
    void* ps[] = {s_cen, s_ord2, s_cb, s_bins, s_splits, s_nodes, s_next, s_small, s_chunks, s_node_chunks, s_chunk_left,
                  s_chunk_base, s_ctr};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (s_tail) (void)hipHostFree(s_tail);
    if (s_meshes) (void)hipFree(s_meshes);
    s_meshes = nullptr;
    s_mesh_cap = 0;
    s_cen = nullptr;
    s_ord2 = s_cb = s_bins = s_chunk_left = s_chunk_base = s_ctr = s_tail = nullptr;
    s_nodes = s_next = s_small = s_chunks = nullptr;
    s_node_chunks = nullptr;
    s_splits = nullptr;
    s_cap = s_nmax = s_cmax = 0;
}

// Binned-SAH binary trees of several meshes at once on the GPU, each in the
// radix tree's form: mesh g = (P, n) owns primitives / positions P..P+n-1
// of d_lo / d_hi / d_order (d_order receives primitive indices) and writes
// its tree to childs + P, range + P, parent + 2P in its own numbering.
static hipError_t sah_forest_gpu(hipStream_t st, const std::vector<int2>& meshes, uint32_t total, const float4* d_lo,
                                 const float4* d_hi, uint32_t* d_order, int2* childs, int2* range, uint32_t* parent,
                                 BuildScratch& sc) {
    HH_TRY(sc.ensure_sah(total));
    HH_TRY(sc.ensure_sah_meshes((uint32_t)meshes.size()));
    const uint32_t B = 256;
    SahSplit* splits = (SahSplit*)sc.s_splits;
    k_sah_centroids<<<blocks(total, B), B, 0, st>>>(d_lo, d_hi, total, sc.s_cen, d_order);
    std::vector<int4> big, small0, chunks;
    std::vector<int2> node_chunks;
    for (size_t g = 0; g < meshes.size(); g++) {
        const int4 root = make_int4(meshes[g].x, meshes[g].x + meshes[g].y, 0, (int)g);
        if (meshes[g].y > kSahSmall) big.push_back(root);
        else if (meshes[g].y >= 2) small0.push_back(root);
    }
    uint32_t ctr0[4] = {0u, (uint32_t)small0.size(), 0u, 0u};   // [0] next big ranges, [1] small ranges
    HH_TRY(hipMemcpyAsync(sc.s_meshes, meshes.data(), meshes.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    HH_TRY(hipMemcpyAsync(sc.s_ctr, ctr0, sizeof(ctr0), hipMemcpyHostToDevice, st));
    if (!small0.empty())
        HH_TRY(hipMemcpyAsync(sc.s_small, small0.data(), small0.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    HH_TRY(hipStreamSynchronize(st));   // the host arrays above are reused
    while (!big.empty()) {
        const uint32_t nn = (uint32_t)big.size();
        if (nn > sc.s_nmax) return hipErrorInvalidValue;
        chunks.clear();
        node_chunks.clear();
        for (uint32_t j = 0; j < nn; j++) {
            const int first = (int)chunks.size();
            for (int s0 = big[j].x; s0 < big[j].y; s0 += kSahChunk)
                chunks.push_back(make_int4((int)j, s0, std::min(big[j].y, s0 + kSahChunk), 0));
            node_chunks.push_back(make_int2(first, (int)chunks.size() - first));
        }
        const uint32_t nc = (uint32_t)chunks.size();
        if (nc > sc.s_cmax) return hipErrorInvalidValue;
        HH_TRY(hipMemcpyAsync(sc.s_nodes, big.data(), nn * sizeof(int4), hipMemcpyHostToDevice, st));
        HH_TRY(hipMemcpyAsync(sc.s_chunks, chunks.data(), nc * sizeof(int4), hipMemcpyHostToDevice, st));
        HH_TRY(hipMemcpyAsync(sc.s_node_chunks, node_chunks.data(), nn * sizeof(int2), hipMemcpyHostToDevice, st));
        k_sah_level_init<<<blocks(nn * kSahBinWords, B), B, 0, st>>>(nn, sc.s_cb, sc.s_bins);
        k_sah_bounds<<<nc, 256, 0, st>>>(sc.s_chunks, d_order, sc.s_cen, sc.s_cb);
        k_sah_bin<<<nc, 256, 0, st>>>(sc.s_chunks, d_order, sc.s_cen, d_lo, d_hi, sc.s_cb, sc.s_bins);
        k_sah_choose<<<nn, 64, 0, st>>>(sc.s_nodes, sc.s_cb, sc.s_bins, splits);
        k_sah_count<<<nc, 256, 0, st>>>(sc.s_chunks, splits, d_order, sc.s_cen, sc.s_chunk_left);
        k_sah_chunkscan<<<blocks(nn, 64), 64, 0, st>>>(nn, sc.s_node_chunks, sc.s_chunk_left, sc.s_chunk_base);
        k_sah_scatter<<<nc, 256, 0, st>>>(sc.s_chunks, sc.s_nodes, splits, d_order, sc.s_cen, sc.s_chunk_base, sc.s_ord2);
        k_sah_copyback<<<nc, 256, 0, st>>>(sc.s_chunks, sc.s_ord2, d_order);
        k_fill_u32<<<1, 64, 0, st>>>(sc.s_ctr, 0u, 1);
        k_sah_emit<<<blocks(nn, 64), 64, 0, st>>>(nn, sc.s_nodes, splits, sc.s_meshes, childs, range, parent, sc.s_next,
                                                 sc.s_ctr, sc.s_small, sc.s_ctr + 1);
        HH_TRY(hipMemcpyAsync(sc.s_tail, sc.s_ctr, 8, hipMemcpyDeviceToHost, st));
        HH_TRY(hipStreamSynchronize(st));
        const uint32_t nnext = sc.s_tail[0];
        if (nnext > sc.s_nmax) return hipErrorInvalidValue;
        big.resize(nnext);
        if (nnext) {
            HH_TRY(hipMemcpyAsync(big.data(), sc.s_next, nnext * sizeof(int4), hipMemcpyDeviceToHost, st));
            HH_TRY(hipStreamSynchronize(st));
        }
    }
    HH_TRY(hipMemcpyAsync(sc.s_tail, sc.s_ctr + 1, 4, hipMemcpyDeviceToHost, st));
    HH_TRY(hipStreamSynchronize(st));
    const uint32_t nsmall = sc.s_tail[0];
    if (nsmall > total / 2 + 1) return hipErrorInvalidValue;
    if (nsmall)
        k_sah_small<<<nsmall, 64, 0, st>>>(sc.s_small, sc.s_meshes, sc.s_cen, d_lo, d_hi, d_order, childs, range, parent);
    return hipGetLastError();
}

hipError_t BuildScratch::ensure_ploc(uint32_t n) {
    if (n <= p_cap) return hipSuccess;
    release_ploc();
    p_cap = std::max<uint32_t>(n, 1024);
    size_t N = p_cap;
    HH_TRY(hipMalloc(&p_sorted, N * 4));
    HH_TRY(hipMalloc(&p_cl[0], N * 4));
    HH_TRY(hipMalloc(&p_cl[1], N * 4));
    HH_TRY(hipMalloc(&p_nn, N * 4));
    HH_TRY(hipMalloc(&p_flags, N * 8));
    HH_TRY(hipMalloc(&p_scan, N * 8));
    HH_TRY(hipMalloc(&p_child, N * sizeof(int2)));
    HH_TRY(hipMalloc(&p_parent, 2 * N * 4));
    HH_TRY(hipMalloc(&p_count, 2 * N * 4));
    HH_TRY(hipMalloc(&p_off, 2 * N * 4));
    HH_TRY(hipHostMalloc(&p_tail, 16));
    scan_bytes = 0;
    HH_TRY(hipcub::DeviceScan::ExclusiveSum((void*)nullptr, scan_bytes, p_flags, p_scan, (int)N));
    HH_TRY(hipMalloc(&scan_tmp, scan_bytes + 256));
    return hipSuccess;
}

void BuildScratch::release_ploc() {
    void* ps[] = {p_sorted, p_cl[0], p_cl[1], p_nn, p_flags, p_scan, p_child, p_parent, p_count, p_off, scan_tmp};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (p_tail) (void)hipHostFree(p_tail);
    p_sorted = p_cl[0] = p_cl[1] = p_nn = p_parent = p_count = p_off = nullptr;
    p_flags = p_scan = p_tail = nullptr;
    p_child = nullptr;
    scan_tmp = nullptr;
    scan_bytes = 0;
    p_cap = 0;
}

hipError_t BuildScratch::ensure(uint32_t n) {
    if (n <= cap) return hipSuccess;
    release();
    cap = std::max<uint32_t>(n, 1024);
    size_t N = cap;
    HH_TRY(hipMalloc(&keys_in, N * 4));
    HH_TRY(hipMalloc(&keys_out, N * 4));
    HH_TRY(hipMalloc(&vals_in, N * 4));
    HH_TRY(hipMalloc(&order, N * 4));
    HH_TRY(hipMalloc(&childs, N * sizeof(int2)));
    HH_TRY(hipMalloc(&range, N * sizeof(int2)));
    HH_TRY(hipMalloc(&parent, 2 * N * 4));
    HH_TRY(hipMalloc(&flags, N * 4));
    HH_TRY(hipMalloc(&blo, 2 * N * sizeof(float4)));
    HH_TRY(hipMalloc(&bhi, 2 * N * sizeof(float4)));
    HH_TRY(hipMalloc(&jobs_a, N * sizeof(int2)));
    HH_TRY(hipMalloc(&jobs_b, N * sizeof(int2)));
    HH_TRY(hipMalloc(&counters, 16));
    HH_TRY(hipMalloc(&cbounds, 32));
    HH_TRY(hipMalloc(&aabb_lo, N * sizeof(float4)));
    HH_TRY(hipMalloc(&aabb_hi, N * sizeof(float4)));
    sort_bytes = 0;
    HH_TRY(hipcub::DeviceRadixSort::SortPairs((void*)nullptr, sort_bytes, keys_in, keys_out, vals_in, vals_in, (int)N, 0, 30));
    HH_TRY(hipMalloc(&sort_tmp, sort_bytes + 256));
    return hipSuccess;
}

void BuildScratch::release() {
    void* ps[] = {keys_in, keys_out, vals_in, order, childs, range, parent, flags, blo, bhi, jobs_a, jobs_b, counters, cbounds,
                  aabb_lo, aabb_hi, sort_tmp};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    keys_in = keys_out = vals_in = order = parent = flags = counters = cbounds = nullptr;
    childs = range = jobs_a = jobs_b = nullptr;
    blo = bhi = aabb_lo = aabb_hi = nullptr;
    sort_tmp = nullptr;
    cap = 0;
}

// BLAS over a triangle soup (3 float4 per triangle, see gpu_scene.hpp):
// writes the leaf-ordered triangles to d_tri_pool + tri_base and the nodes to
// d_node_pool + node_base (global indices in child references).
hipError_t build_blas(hipStream_t st, int width, int leaf_size, uint32_t n, const float4* d_tri_src, float4* d_tri_pool,
                      const float4* d_shade_src, float4* d_shade_pool, const float4* h_tri_src, uint32_t tri_base, void* d_node_pool, uint32_t node_base, BuildScratch& sc, uint32_t* num_nodes,
                      float root_box[6], bool ploc, bool sah_gpu) {
    HH_TRY(sc.ensure(n));
    const uint32_t B = 256;
    k_init_bounds<<<1, 64, 0, st>>>(sc.cbounds);
    k_tri_aabbs<<<blocks(n, B), B, 0, st>>>(d_tri_src, n, sc.aabb_lo, sc.aabb_hi, sc.cbounds);
    uint32_t* order = sc.vals_out_tmp(n);
    std::vector<float> hlo, hhi;   // host triangle boxes for the SAH build
    if (h_tri_src) {
        hlo.resize(size_t(n) * 3);
        hhi.resize(size_t(n) * 3);
        for (uint32_t t = 0; t < n; t++) {
            const float4 a = h_tri_src[3 * t], b = h_tri_src[3 * t + 1], c = h_tri_src[3 * t + 2];
            const float va[3] = {a.x, a.y, a.z}, vb[3] = {b.x, b.y, b.z}, vc[3] = {c.x, c.y, c.z};
            for (int k = 0; k < 3; k++) {
                hlo[3 * t + k] = std::min(va[k], std::min(vb[k], vc[k]));
                hhi[3 * t + k] = std::max(va[k], std::max(vb[k], vc[k]));
            }
        }
    }
    const float* pl = h_tri_src ? hlo.data() : nullptr;
    const float* ph = h_tri_src ? hhi.data() : nullptr;
    hipError_t e;
    if (width == 8)
        e = build_wide<8>(st, n, sc.aabb_lo, sc.aabb_hi, sc.cbounds, (WideNode<8>*)d_node_pool + node_base, node_base,
                          tri_base, leaf_size, 0, order, sc, num_nodes, root_box, pl, ph, ploc, sah_gpu);
    else
        e = build_wide<4>(st, n, sc.aabb_lo, sc.aabb_hi, sc.cbounds, (WideNode<4>*)d_node_pool + node_base, node_base,
                          tri_base, leaf_size, 0, order, sc, num_nodes, root_box, pl, ph, ploc, sah_gpu);
    HH_TRY(e);
    k_gather_tris<<<blocks(n, B), B, 0, st>>>(d_tri_src, order, n, d_tri_pool + 3 * (size_t)tri_base);
    k_gather_shade<<<blocks(n, B), B, 0, st>>>(d_shade_src, order, n, d_shade_pool + 4 * (size_t)tri_base);
    return hipGetLastError();
}

// BLASes of several meshes over one triangle soup, by the GPU binned-SAH
// builder: mesh g = (P, n) owns triangles P..P+n-1 of d_tri_src and of the
// pools. The trees of all meshes are built together (one pass per level
// over every mesh's large ranges, one launch for all small ranges); then
// per mesh the bounds and the wide collapse, whose nodes start at
// node_base (advancing mesh by mesh: roots[g], counts[g]); then the
// triangles and shading records of all meshes are gathered into leaf order.
hipError_t build_blas_forest(hipStream_t st, int width, int leaf_size, const std::vector<int2>& meshes, uint32_t total,
                             const float4* d_tri_src, float4* d_tri_pool, const float4* d_shade_src, float4* d_shade_pool,
                             void* d_node_pool, uint32_t node_base, BuildScratch& sc, uint32_t* roots, uint32_t* counts,
                             float* root_boxes) {
    if (meshes.empty() || total == 0) return hipSuccess;
    HH_TRY(sc.ensure(total));
    const uint32_t B = 256;
    k_init_bounds<<<1, 64, 0, st>>>(sc.cbounds);
    k_tri_aabbs<<<blocks(total, B), B, 0, st>>>(d_tri_src, total, sc.aabb_lo, sc.aabb_hi, sc.cbounds);
    HH_TRY(sah_forest_gpu(st, meshes, total, sc.aabb_lo, sc.aabb_hi, sc.order, sc.childs, sc.range, sc.parent, sc));
    k_fill_u32<<<blocks(total, B), B, 0, st>>>(sc.flags, 0u, total);
    for (size_t g = 0; g < meshes.size(); g++) {
        const uint32_t P = (uint32_t)meshes[g].x, n = (uint32_t)meshes[g].y;
        uint32_t nn = 0;
        hipError_t e;
        if (width == 8)
            e = finish_wide<8>(st, n, sc.aabb_lo, sc.aabb_hi, sc.order + P, sc.childs + P, sc.range + P, sc.parent + 2 * (size_t)P,
                               sc.flags + P, sc.blo + 2 * (size_t)P, sc.bhi + 2 * (size_t)P, sc, leaf_size, 0,
                               (WideNode<8>*)d_node_pool + node_base, node_base, P, &nn, root_boxes + 6 * g);
        else
            e = finish_wide<4>(st, n, sc.aabb_lo, sc.aabb_hi, sc.order + P, sc.childs + P, sc.range + P, sc.parent + 2 * (size_t)P,
                               sc.flags + P, sc.blo + 2 * (size_t)P, sc.bhi + 2 * (size_t)P, sc, leaf_size, 0,
                               (WideNode<4>*)d_node_pool + node_base, node_base, P, &nn, root_boxes + 6 * g);
        HH_TRY(e);
        roots[g] = node_base;
        counts[g] = nn;
        node_base += nn;
    }
    k_gather_tris<<<blocks(total, B), B, 0, st>>>(d_tri_src, sc.order, total, d_tri_pool);
    k_gather_shade<<<blocks(total, B), B, 0, st>>>(d_shade_src, sc.order, total, d_shade_pool);
    return hipGetLastError();
}

hipError_t quantize_bvh4(hipStream_t st, const void* d_src, void* d_dst, uint32_t n) {
    if (n == 0) return hipSuccess;
    k_quantize4<<<blocks(n, 128), 128, 0, st>>>((const WideNode<4>*)d_src, (QNode4*)d_dst, n);
    return hipGetLastError();
}

// TLAS over instance AABBs (host arrays, n >= 1); leaves hold one instance.
hipError_t build_tlas(hipStream_t st, int width, uint32_t n, const float* h_lo, const float* h_hi, void* d_node_pool,
                      uint32_t node_base, BuildScratch& sc, uint32_t* num_nodes, bool sah, bool ploc, bool sah_gpu) {
    HH_TRY(sc.ensure(n));
    std::vector<float4> lo(n), hi(n);
    for (uint32_t i = 0; i < n; i++) {
        lo[i] = make_float4(h_lo[3 * i], h_lo[3 * i + 1], h_lo[3 * i + 2], 0.f);
        hi[i] = make_float4(h_hi[3 * i], h_hi[3 * i + 1], h_hi[3 * i + 2], 0.f);
    }
    HH_TRY(hipMemcpyAsync(sc.aabb_lo, lo.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
    HH_TRY(hipMemcpyAsync(sc.aabb_hi, hi.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
    const uint32_t B = 256;
    k_init_bounds<<<1, 64, 0, st>>>(sc.cbounds);
    k_box_centroids<<<blocks(n, B), B, 0, st>>>(sc.aabb_lo, sc.aabb_hi, n, sc.cbounds);
    uint32_t* order = sc.vals_out_tmp(n);
    float rb[6];
    const float* pl = sah ? h_lo : nullptr;
    const float* ph = sah ? h_hi : nullptr;
    if (width == 8)
        return build_wide<8>(st, n, sc.aabb_lo, sc.aabb_hi, sc.cbounds, (WideNode<8>*)d_node_pool + node_base, node_base, 0, 1,
                             1, order, sc, num_nodes, rb, pl, ph, ploc, sah_gpu);
    return build_wide<4>(st, n, sc.aabb_lo, sc.aabb_hi, sc.cbounds, (WideNode<4>*)d_node_pool + node_base, node_base, 0, 1, 1,
                         order, sc, num_nodes, rb, pl, ph, ploc, sah_gpu);
}

}  // namespace hh
