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
    WideNode<W> nd;
    for (int k = 0; k < W; k++) {
        nd.lox[k] = nd.loy[k] = nd.loz[k] = 0.f;
        nd.hix[k] = nd.hiy[k] = nd.hiz[k] = 0.f;
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
// surface-area cost (leaf of m primitives costs m, a split costs its
// children's area-weighted counts); ranges whose centroids coincide, or
// whose best split would leave a side empty, are split at the median. The
// subtree over m primitives owns the m - 1 internal ids after its root, so
// subtrees are built in parallel without shared state.
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

    static float area(const float lo[3], const float hi[3]) {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return dx * dy + dy * dz + dz * dx;
    }

    // split [b, e) and return the first index of the right side
    int split(int b, int e) {
        const std::vector<SahPrim>& P = *prims;
        float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = b; i < e; i++)
            for (int k = 0; k < 3; k++) {
                cl[k] = std::min(cl[k], P[order[i]].c[k]);
                ch[k] = std::max(ch[k], P[order[i]].c[k]);
            }
        constexpr int NB = 32;
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
                const int q = std::min(NB - 1, int((p.c[k] - cl[k]) * scale));
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
                const float cost = lc * area(ll, lh) + rcnt[q + 1] * rarea[q + 1];
                if (cost < best) { best = cost; best_axis = k; best_bin = q; }
            }
        }
        if (best_axis >= 0) {
            const float ext = ch[best_axis] - cl[best_axis], scale = NB / ext;
            auto mid = std::partition(order.begin() + b, order.begin() + e, [&](uint32_t i) {
                return std::min(NB - 1, int((P[i].c[best_axis] - cl[best_axis]) * scale)) <= best_bin;
            });
            const int m = int(mid - order.begin());
            if (m > b && m < e) return m;
        }
        // coincident centroids or a one-sided split: median on the widest axis
        int ax = 0;
        for (int k = 1; k < 3; k++)
            if (ch[k] - cl[k] > ch[ax] - cl[ax]) ax = k;
        const int m = b + (e - b) / 2;
        std::nth_element(order.begin() + b, order.begin() + m, order.begin() + e,
                         [&](uint32_t x, uint32_t y) { return P[x].c[ax] < P[y].c[ax]; });
        return m;
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

template <int W>
static hipError_t build_wide(hipStream_t st, uint32_t n, const float4* d_lo, const float4* d_hi, const uint32_t* d_cbounds,
                             WideNode<W>* d_nodes, uint32_t node_base, uint32_t leaf_base, int leaf_size, int leaf_is_prim,
                             uint32_t* d_order, BuildScratch& sc, uint32_t* num_nodes, float root_box[6],
                             const float* h_lo = nullptr, const float* h_hi = nullptr, bool ploc = false) {
    HH_TRY(sc.ensure(n));
    const uint32_t B = 256;
    if (ploc && n > 1) {
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
    k_refit<<<blocks(n, B), B, 0, st>>>(d_lo, d_hi, d_order, (int)n, sc.childs, sc.parent, sc.flags, sc.blo, sc.bhi);
    int root = n > 1 ? 0 : 0;  // binary ref of the root (leaf 0 when n == 1: ref n-1+0 = 0)
    int2 job0 = make_int2(root, 0);
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
        k_collapse<W><<<blocks(njobs, 64), 64, 0, st>>>(qa, ca, qb, cb, sc.counters + 2, (int)n, leaf_size, sc.childs, sc.range,
                                                         sc.blo, sc.bhi, d_order, d_nodes, node_base, leaf_base, leaf_is_prim);
        HH_TRY(hipMemcpyAsync(&njobs, cb, 4, hipMemcpyDeviceToHost, st));
        HH_TRY(hipStreamSynchronize(st));
        std::swap(qa, qb);
        std::swap(ca, cb);
    }
    HH_TRY(hipMemcpyAsync(num_nodes, sc.counters + 2, 4, hipMemcpyDeviceToHost, st));
    float4 rb[2];
    HH_TRY(hipMemcpyAsync(&rb[0], &sc.blo[0], 16, hipMemcpyDeviceToHost, st));
    HH_TRY(hipMemcpyAsync(&rb[1], &sc.bhi[0], 16, hipMemcpyDeviceToHost, st));
    HH_TRY(hipStreamSynchronize(st));
    root_box[0] = rb[0].x; root_box[1] = rb[0].y; root_box[2] = rb[0].z;
    root_box[3] = rb[1].x; root_box[4] = rb[1].y; root_box[5] = rb[1].z;
    return hipSuccess;
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
                      float root_box[6], bool ploc) {
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
                          tri_base, leaf_size, 0, order, sc, num_nodes, root_box, pl, ph, ploc);
    else
        e = build_wide<4>(st, n, sc.aabb_lo, sc.aabb_hi, sc.cbounds, (WideNode<4>*)d_node_pool + node_base, node_base,
                          tri_base, leaf_size, 0, order, sc, num_nodes, root_box, pl, ph, ploc);
    HH_TRY(e);
    k_gather_tris<<<blocks(n, B), B, 0, st>>>(d_tri_src, order, n, d_tri_pool + 3 * (size_t)tri_base);
    k_gather_shade<<<blocks(n, B), B, 0, st>>>(d_shade_src, order, n, d_shade_pool + 4 * (size_t)tri_base);
    return hipGetLastError();
}

hipError_t quantize_bvh4(hipStream_t st, const void* d_src, void* d_dst, uint32_t n) {
    if (n == 0) return hipSuccess;
    k_quantize4<<<blocks(n, 128), 128, 0, st>>>((const WideNode<4>*)d_src, (QNode4*)d_dst, n);
    return hipGetLastError();
}

// TLAS over instance AABBs (host arrays, n >= 1); leaves hold one instance.
hipError_t build_tlas(hipStream_t st, int width, uint32_t n, const float* h_lo, const float* h_hi, void* d_node_pool,
                      uint32_t node_base, BuildScratch& sc, uint32_t* num_nodes, bool sah, bool ploc) {
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
                             1, order, sc, num_nodes, rb, pl, ph, ploc);
    return build_wide<4>(st, n, sc.aabb_lo, sc.aabb_hi, sc.cbounds, (WideNode<4>*)d_node_pool + node_base, node_base, 0, 1, 1,
                         order, sc, num_nodes, rb, pl, ph, ploc);
}

}  // namespace hh
