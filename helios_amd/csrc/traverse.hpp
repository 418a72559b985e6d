// traverse.hpp — two-level wide-BVH traversal with watertight triangle tests.
//
// Replaces the closest-hit / any-hit / terminate-on-first-hit semantics of
// traceRayEXT over the instance TLAS and per-mesh BLAS
// (src/engine/shader/path_trace_rgen.glsl:205-215,
//  src/engine/shader/path_trace_rchit.glsl:438-448,523-533; any-hit
//  alpha test path_trace_rahit.glsl:174-188; BLAS geometry opaque flag
//  src/engine/resource/mesh.cpp:75-76; instance cull disabled
//  src/engine/resource/scene.cpp:1257).
//
// Per lane: a short stack of kLdsStack entries in LDS (interleaved by lane so
// the 64 lanes of a wave hit 64 different banks) that spills into a global
// per-thread area when deeper (no private/scratch memory). Boxes use the robust slab test of Ize ("Robust BVH Ray
// Traversal", JCGT 2013); triangles use the watertight test of Woop, Benthin
// and Wald (JCGT 2013) with its double-precision fallback on zero edge
// functions. Equal-t hits are resolved to the smallest (instance, geometry,
// primitive), so the result does not depend on traversal order.
#pragma once

#include "gpu_common.hpp"
#include "gpu_scene.hpp"

namespace hh {

constexpr int kLdsStack = 24;
constexpr int kSpillStack = kSpillStackDepth;

// Explicit address spaces: pointers that arrive through structs are generic,
// and generic (flat) accesses both serialize on two wait counters and, for
// the LDS stack, bypass the ds_* path. Traversal casts to these once. The
// host pass only parses device bodies, so there the qualifiers are empty.
#if defined(__HIP_DEVICE_COMPILE__)
#define HH_GLOBAL __attribute__((address_space(1)))
#define HH_LDS __attribute__((address_space(3)))
#else
#define HH_GLOBAL
#define HH_LDS
#endif

// Where one lane keeps its traversal stack: the first kLdsStack entries in LDS
// (entry k at lds[k * lds_stride]), deeper entries in a global per-thread
// spill area (entry k at spill[(k - kLdsStack) * spill_stride]). Keeping the
// spill out of private memory keeps the kernels free of scratch.
struct LaneStack {
    HH_LDS uint32_t* lds;
    uint32_t lds_stride;
    HH_GLOBAL uint32_t* spill;
    uint32_t spill_stride;
};
__device__ inline HH_LDS uint32_t* to_lds(uint32_t* p) { return (HH_LDS uint32_t*)p; }
__device__ inline HH_GLOBAL uint32_t* to_global(uint32_t* p) { return (HH_GLOBAL uint32_t*)p; }

// Per-ray fetch counts of the counting variant (algorithmic bytes).
struct TravCount {
    uint32_t nodes, tris, insts;
};

struct HitRec {
    float t, u, v;
    uint32_t inst;
    uint32_t tri;  // global triangle index
    bool found;
};

struct WT {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
__device__ inline float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ inline WT wt_pre(f3 d) {
    WT p;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    p.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    p.kx = p.kz + 1;
    if (p.kx == 3) p.kx = 0;
    p.ky = p.kx + 1;
    if (p.ky == 3) p.ky = 0;
    if (comp(d, p.kz) < 0.0f) {
        int t = p.kx;
        p.kx = p.ky;
        p.ky = t;
    }
    float dz = comp(d, p.kz);
    p.Sx = comp(d, p.kx) / dz;
    p.Sy = comp(d, p.ky) / dz;
    p.Sz = 1.0f / dz;
    return p;
}

// Woop-Benthin-Wald watertight test; the operation order is fixed (the
// library is compiled without FP contraction) and equals the CPU oracle's.
__device__ inline bool tri_hit(const WT& pr, f3 o, f3 v0, f3 v1, f3 v2, float tmin, float tmax, float& t, float& u,
                               float& v) {
    f3 A = v0 - o, B = v1 - o, C = v2 - o;
    float Akz = comp(A, pr.kz), Bkz = comp(B, pr.kz), Ckz = comp(C, pr.kz);
    float Ax = comp(A, pr.kx) - pr.Sx * Akz, Ay = comp(A, pr.ky) - pr.Sy * Akz;
    float Bx = comp(B, pr.kx) - pr.Sx * Bkz, By = comp(B, pr.ky) - pr.Sy * Bkz;
    float Cx = comp(C, pr.kx) - pr.Sx * Ckz, Cy = comp(C, pr.ky) - pr.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = pr.Sz * Akz, Bz = pr.Sz * Bkz, Cz = pr.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    u = V / det;
    v = W / det;
    return true;
}

// World -> object (row-major 3x4), fixed operation order.
template <typename P>
__device__ inline void xform_ray(P m, f3 o, f3 d, f3& oo, f3& od) {
    oo.x = ((m[0] * o.x + m[1] * o.y) + m[2] * o.z) + m[3];
    oo.y = ((m[4] * o.x + m[5] * o.y) + m[6] * o.z) + m[7];
    oo.z = ((m[8] * o.x + m[9] * o.y) + m[10] * o.z) + m[11];
    od.x = (m[0] * d.x + m[1] * d.y) + m[2] * d.z;
    od.y = (m[4] * d.x + m[5] * d.y) + m[6] * d.z;
    od.z = (m[8] * d.x + m[9] * d.y) + m[10] * d.z;
}

// ---------------------------------------------------------------------------
// Texture sampling (Vulkan bilinear, LOD 0, REPEAT) and the any-hit alpha test.
// ---------------------------------------------------------------------------
__device__ inline float lerpf(float a, float b, float w) { return a + (b - a) * w; }
__device__ inline int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}
__device__ inline float4 sample_tex(const SceneGPU& S, int slot, float s, float t) {
    TexDesc td = S.tex[slot];
    float x = s * (float)td.w - 0.5f, y = t * (float)td.h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int i0 = (int)fx, j0 = (int)fy;
    int i1 = wrapi(i0 + 1, (int)td.w), j1 = wrapi(j0 + 1, (int)td.h);
    i0 = wrapi(i0, (int)td.w);
    j0 = wrapi(j0, (int)td.h);
    const float4* base = S.texels + td.offset;
    float4 p00 = base[j0 * td.w + i0], p10 = base[j0 * td.w + i1];
    float4 p01 = base[j1 * td.w + i0], p11 = base[j1 * td.w + i1];
    float4 r;
    r.x = lerpf(lerpf(p00.x, p10.x, ax), lerpf(p01.x, p11.x, ax), ay);
    r.y = lerpf(lerpf(p00.y, p10.y, ax), lerpf(p01.y, p11.y, ax), ay);
    r.z = lerpf(lerpf(p00.z, p10.z, ax), lerpf(p01.z, p11.z, ax), ay);
    r.w = lerpf(lerpf(p00.w, p10.w, ax), lerpf(p01.w, p11.w, ax), ay);
    return r;
}

// path_trace_rahit.glsl:162-188: alpha of the candidate hit on triangle
// `tri` (global index; its UVs come from the shading record).
// The instance record names the alpha source of its non-opaque geometry when
// it is the same for all of it, which saves the submesh and material lookups.
__device__ inline float hit_alpha(const SceneGPU& S, uint32_t inst, uint32_t geom, uint32_t tri, float u, float v) {
    const uint2 src = *reinterpret_cast<const uint2*>(&S.inst_trav[inst].alpha_src);
    const float4 r0 = S.shade[4 * (size_t)tri], r1 = S.shade[4 * (size_t)tri + 1];
    int slot = (int)src.x;
    if (src.x == kAlphaConst) return __uint_as_float(src.y);
    if (src.x == kAlphaMixed) {
        const uint2 info = S.submesh_info[S.submesh_off[inst] + geom];
        const MaterialData& m = S.mats[info.y];
        if (m.tex0[0] == -1) return m.albedo[3];
        slot = m.tex0[0];
    }
    float bx = (1.0f - u) - v, by = u, bz = v;
    float tu = (r0.x * bx + r0.z * by) + r1.x * bz;
    float tv = (r0.y * bx + r0.w * by) + r1.y * bz;
    return sample_tex(S, slot, tu, tv).w;
}

// Reciprocal direction for the box tests: v_rcp_f32 (within 1 ulp of 1/d,
// +-inf for +-0) instead of a correctly rounded division (about ten
// instructions each). The triangle test does not use it.
__device__ inline f3 box_inv(f3 d) {
    return mk(__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z));
}

// Padding of the far distance in the slab test. Each slab distance
// fl(fl(lo - o) * rcp(d)) carries at most 4u of relative error (u = 2^-24:
// one rounding each for the subtraction and the product, 2u for the
// reciprocal), so a computed entry exceeds the computed exit of a box the
// exact ray meets by at most (1 + 4u) / (1 - 4u), and the product with the
// padding rounds by one more u: 1 + 10u covers it (Ize's gamma_n argument with
// n = 4), so a box is never culled that the exact ray enters.
constexpr float kBoxPad = 0x1.00000ap+0f;

// One slab test in the robust form (Ize 2013): returns the entry distance or
// +inf on a miss. min/max drop a NaN operand (0 * inf for zero directions).
__device__ inline float slab(float lox, float hix, float loy, float hiy, float loz, float hiz, f3 o, f3 inv, float tmin,
                             float tmax) {
    float t0x = (lox - o.x) * inv.x, t1x = (hix - o.x) * inv.x;
    float t0y = (loy - o.y) * inv.y, t1y = (hiy - o.y) * inv.y;
    float t0z = (loz - o.z) * inv.z, t1z = (hiz - o.z) * inv.z;
    float a = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    float b = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    return a <= b * kBoxPad ? a : INFINITY;
}

__device__ inline float exp2i(int e) { return __int_as_float((e + 127) << 23); }
__device__ inline float qbyte(uint32_t w, int k) { return (float)((w >> (8 * k)) & 255u); }

// Tests the children of node `cur`: tn[k] = entry distance (+inf = miss or
// absent child), ref[k] = child reference.
template <int L>
__device__ inline void node_test(const HH_GLOBAL void* nodes, uint32_t cur, f3 o, f3 inv, float tmin, float tmax, float* tn,
                                 uint32_t* ref) {
    if constexpr (L == kLayoutQ4) {
        const HH_GLOBAL uint4* p = (const HH_GLOBAL uint4*)nodes + 4 * (size_t)cur;
        const uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3];
        const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
        const float sx = exp2i((int)(int8_t)(w0.w & 255u)), sy = exp2i((int)(int8_t)((w0.w >> 8) & 255u)),
                    sz = exp2i((int)(int8_t)((w0.w >> 16) & 255u));
        const uint32_t nc = w0.w >> 24;
        ref[0] = w2.z;
        ref[1] = w2.w;
        ref[2] = w3.x;
        ref[3] = w3.y;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float t = slab(fmaf(qbyte(w1.x, k), sx, ox), fmaf(qbyte(w1.y, k), sx, ox), fmaf(qbyte(w1.z, k), sy, oy),
                           fmaf(qbyte(w1.w, k), sy, oy), fmaf(qbyte(w2.x, k), sz, oz), fmaf(qbyte(w2.y, k), sz, oz), o, inv,
                           tmin, tmax);
            tn[k] = (uint32_t)k < nc ? t : INFINITY;
        }
    } else {
        constexpr int W = layout_width(L);
        // absent children have all-infinite boxes, which every ray misses
        const HH_GLOBAL WideNode<W>& nd = ((const HH_GLOBAL WideNode<W>*)nodes)[cur];
#pragma unroll
        for (int k = 0; k < W; k++) {
            tn[k] = slab(nd.lox[k], nd.hix[k], nd.loy[k], nd.hiy[k], nd.loz[k], nd.hiz[k], o, inv, tmin, tmax);
            ref[k] = nd.child[k];
        }
    }
}

__device__ inline bool key_less(uint32_t i1, uint32_t g1, uint32_t p1, uint32_t i2, uint32_t g2, uint32_t p2) {
    if (i1 != i2) return i1 < i2;
    if (g1 != g2) return g1 < g2;
    return p1 < p2;
}

// ---------------------------------------------------------------------------
// Traversal as a per-lane state machine. ANY = stop at the first accepted hit
// (visibility rays). rflags: HH_TRACE_OPAQUE (bit 0) disables the any-hit
// alpha test. `step` advances one lane; the persistent queue kernels interleave
// steps with refills of finished lanes, so a wave never idles on the longest
// ray of a batch. The result (closest (t, instance, geometry, primitive)) does
// not depend on the visiting order, so every schedule is bit-identical.
// ---------------------------------------------------------------------------
constexpr uint32_t kDone = 0xFFFFFFFEu;     // never a valid child reference
constexpr uint32_t kMaxTrips = 1u << 22;    // hard bound on node + leaf visits per ray
constexpr uint32_t kStackFull = 0xFFFFFFFFu;  // trip count of a ray whose stack overflowed

// Step shapes: if-if visits one node or leaf per step; while-while visits
// inner nodes until the lane reaches a leaf (or finishes), then one leaf.
constexpr uint32_t kLeafWait = 3;
enum TravMode { kWhileWhile = 0, kIfIf = 1 };

template <int L, bool ANY, bool COUNT>
struct Tracer {
    static constexpr int W = layout_width(L);
    const SceneGPU* S;
    // scene arrays read on every visit, held in registers: reading them
    // through S would put a dependent load in front of every node fetch (the
    // instance records, read once per instance entry, are reached through S)
    const HH_GLOBAL void* nodes;
    const HH_GLOBAL float4* tris;
    f3 wo, wd, winv;   // world-space ray
    f3 o, inv;         // ray in the current (world or instance) space
    WT wt;
    float tmin;
    int inst;          // instance being traversed, -1 at TLAS level
    uint32_t sp, cur, trips;
    bool alpha_on;
    HitRec h;
    uint32_t hgeom, hprim;
    TravCount* cnt;

    __device__ inline void begin(const SceneGPU& scene, f3 ro, f3 rd, float t0, float t1, uint32_t rflags,
                                 TravCount* c = nullptr) {
        S = &scene;
        nodes = (const HH_GLOBAL void*)scene.nodes;
        tris = (const HH_GLOBAL float4*)scene.tris;
        wo = ro;
        wd = rd;
        winv = box_inv(rd);
        o = wo;
        inv = winv;
        wt = wt_pre(rd);
        tmin = t0;
        inst = -1;
        sp = 0;
        trips = 0;
        alpha_on = (rflags & 1u) == 0u;
        h.found = false;
        h.t = t1;
        h.u = h.v = 0.0f;
        h.inst = 0xFFFFFFFFu;
        h.tri = 0xFFFFFFFFu;
        hgeom = hprim = 0xFFFFFFFFu;
        cnt = c;
        cur = scene.tlas_root;
    }

    // A push beyond the stack capacity is not stored (never written out of
    // bounds) but still counted in sp; check_full, called once after a
    // visit's pushes and before any pop, then ends the ray: `trips` takes
    // kStackFull, which stops the visit loop, and the scene's fault counter
    // records it. The capacity is far above the depth of the built trees.
    __device__ inline void push(const LaneStack& stk, uint32_t r) {
        if (sp < (uint32_t)kLdsStack) stk.lds[sp * stk.lds_stride] = r;
        else if (sp < (uint32_t)(kLdsStack + kSpillStack)) stk.spill[(sp - kLdsStack) * stk.spill_stride] = r;
        sp++;
    }
    __device__ inline void check_full() {
        if (sp > (uint32_t)(kLdsStack + kSpillStack)) {
            trips = kStackFull;
            sp = kLdsStack + kSpillStack;
        }
    }
    // The LDS entry is read unconditionally (entry 0 when the stack is empty,
    // the value then discarded): only the rare spill read needs a branch.
    __device__ inline void pop(const LaneStack& stk) {
        const uint32_t s = sp == 0u ? 0u : sp - 1u;
        uint32_t r = stk.lds[min(s, (uint32_t)kLdsStack - 1u) * stk.lds_stride];
        if (s >= (uint32_t)kLdsStack) r = stk.spill[(s - kLdsStack) * stk.spill_stride];
        cur = sp == 0u ? kDone : r;
        sp = s;
    }

    // Children of inner node `cur` in front-to-back order: descend into the
    // nearest, push the others, or pop when none is hit.
    __device__ inline void visit_inner(const LaneStack& stk) {
        if (COUNT) cnt->nodes++;
        float tn[W];
        uint32_t ref[W];
        node_test<L>(nodes, cur, o, inv, tmin, h.t, tn, ref);
        // With room for every child in the LDS part of the stack, children are
        // stored without branches: each is written at the next free entry,
        // which advances only when that child is hit (a missed child's entry
        // is overwritten by the next one or left above the top).
        const bool room = sp + (uint32_t)W <= (uint32_t)kLdsStack;
        if (ANY) {
            // any accepted hit ends the ray, so the order is free: children
            // in node order (measured to find occluders in fewer visits
            // than nearest-first on the sun and point-light rays). The
            // lowest hit child is visited next, the others pushed above
            // the stack's top in descending order.
            if (room) {
                uint32_t s = sp, next = kDone;
#pragma unroll
                for (int k = W - 1; k >= 1; k--) {
                    stk.lds[s * stk.lds_stride] = ref[k];
                    const bool hit = tn[k] != INFINITY;
                    s += hit ? 1u : 0u;
                    next = hit ? ref[k] : next;
                }
                if (tn[0] != INFINITY) {
                    cur = ref[0];
                    sp = s;
                } else if (s != sp) {
                    cur = next;
                    sp = s - 1u;
                } else {
                    pop(stk);
                }
                return;
            }
#pragma unroll
            for (int k = W - 1; k >= 0; k--)
                if (tn[k] != INFINITY) push(stk, ref[k]);
            check_full();
            pop(stk);
            return;
        }
        auto cswap = [&](int i, int j) {   // selects, not branches
            const bool c = tn[j] < tn[i];
            const float a = tn[i], b = tn[j];
            const uint32_t ra = ref[i], rb = ref[j];
            tn[i] = c ? b : a;
            tn[j] = c ? a : b;
            ref[i] = c ? rb : ra;
            ref[j] = c ? ra : rb;
        };
        if constexpr (W == 4) {
            // the 5-comparator sorting network for 4 keys
            cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
        } else {
#pragma unroll
            for (int i = 1; i < W; i++)
#pragma unroll
                for (int j = W - 1; j >= i; j--) cswap(j - 1, j);
        }
        if (tn[0] == INFINITY) {
            pop(stk);
            return;
        }
        if (room) {
            uint32_t s = sp;
#pragma unroll
            for (int k = W - 1; k >= 1; k--) {
                stk.lds[s * stk.lds_stride] = ref[k];
                s += tn[k] != INFINITY ? 1u : 0u;
            }
            sp = s;
        } else {
#pragma unroll
            for (int k = W - 1; k >= 1; k--)
                if (tn[k] != INFINITY) push(stk, ref[k]);
            check_full();
        }
        cur = ref[0];
    }

    // Triangles of leaf `cur` (BLAS level); returns true when an any-hit ray
    // found its occluder.
    __device__ inline bool visit_leaf() {
        const uint32_t first = leaf_start(cur), ntri = leaf_count(cur);
        for (uint32_t k = 0; k < ntri; k++) {
            const uint32_t ti = first + k;
            const float4 a = tris[3 * ti], b = tris[3 * ti + 1], c = tris[3 * ti + 2];
            if (COUNT) cnt->tris++;
            float t, u, v;
            if (!tri_hit(wt, o, ld3(a), ld3(b), ld3(c), tmin, h.t, t, u, v)) continue;
            const uint32_t prim = __float_as_uint(a.w);
            const uint32_t gw = __float_as_uint(b.w);
            const uint32_t geom = gw & 0x7FFFFFFFu;
            if (h.found && t == h.t && !key_less((uint32_t)inst, geom, prim, h.inst, hgeom, hprim)) continue;
            if (alpha_on && (gw & 0x80000000u) == 0u) {
                if (hit_alpha(*S, (uint32_t)inst, geom, ti, u, v) < 0.1f) continue;
            }
            h.found = true;
            h.t = t;
            h.u = u;
            h.v = v;
            h.inst = (uint32_t)inst;
            h.tri = ti;
            hgeom = geom;
            hprim = prim;
            if (ANY) return true;
        }
        return false;
    }

    // The lane is at an inner node with visits left: (cur & kLeafBit) == 0 and
    // trips < kMaxTrips = 2^22, tested as one sign compare. trips never
    // exceeds kMaxTrips except as kStackFull, whose shifted value also has
    // bit 31 set.
    __device__ inline bool at_inner() const {
        static_assert(kLeafBit == 0x80000000u && kMaxTrips == (1u << 22), "one-compare inner test");
        return (int32_t)(cur | (trips << 9)) >= 0;
    }

    // Advances the traversal; returns true once the ray is finished (h holds
    // the result).
    template <int MODE>
    __device__ inline bool step(const LaneStack& stk) {
        if (MODE == kWhileWhile) {
            // inner nodes until this lane reaches a leaf, or until the lanes
            // waiting at a leaf are kLeafWait times as many as those still in
            // inner nodes (the wave then visits the waiting lanes' leaves and
            // the others resume their inner nodes at the next step). Live
            // lanes (ml) are inner or waiting, so the loop runs while
            // live < (kLeafWait + 1) * inner. The masks come straight from
            // one compare each and the inner lanes are selected by their
            // mask bit, so no per-lane flag is kept across the loop's blocks.
            unsigned long long mi = __builtin_amdgcn_ballot_w64(at_inner());
            unsigned long long ml = __builtin_amdgcn_ballot_w64(cur != kDone);
            while (mi != 0ull && (uint32_t)__popcll(ml) < (kLeafWait + 1u) * (uint32_t)__popcll(mi)) {
                if (__builtin_amdgcn_inverse_ballot_w64(mi)) {
                    trips++;
                    visit_inner(stk);
                }
                mi = __builtin_amdgcn_ballot_w64(at_inner());
                ml = __builtin_amdgcn_ballot_w64(cur != kDone);
            }
            if (at_inner()) return false;
        } else if (at_inner()) {
            trips++;
            visit_inner(stk);
            return false;
        }
        if (cur == kDone) return true;
        if (trips >= kMaxTrips) {  // visit bound reached, or a full stack
            atomicAdd(&S->faults[trips == kStackFull ? kFaultLaneStack : kFaultTrips], 1ull);
            return true;
        }
        trips++;
        if (cur == kMarker) {
            o = wo;
            inv = winv;
            inst = -1;
            pop(stk);
        } else if (inst < 0) {
            inst = (int)leaf_start(cur);
            const HH_GLOBAL InstTrav& it = ((const HH_GLOBAL InstTrav*)S->inst_trav)[inst];
            if (COUNT) cnt->insts++;
            f3 od;
            xform_ray(it.w2o, wo, wd, o, od);
            inv = box_inv(od);
            wt = wt_pre(od);
            push(stk, kMarker);
            check_full();
            cur = it.blas_root;
            return false;
        } else {
            if (visit_leaf()) return true;
            pop(stk);
        }
        return cur == kDone;
    }
};

// Whole traversal of one ray in one call (debug entry points, megakernel).
template <int L, bool ANY, bool COUNT = false>
__device__ inline bool trace_ray(const SceneGPU& S, f3 wo, f3 wd, float tmin, float tmax, uint32_t rflags, HitRec& h,
                                 const LaneStack& stk, TravCount* cnt = nullptr) {
    Tracer<L, ANY, COUNT> tr;
    tr.begin(S, wo, wd, tmin, tmax, rflags, cnt);
    while (!tr.template step<kWhileWhile>(stk)) {
    }
    h = tr.h;
    return h.found;
}

// ---------------------------------------------------------------------------
// Packet traversal for neighbouring camera rays (the 64 lanes of a wave hold
// jittered samples of one pixel, or neighbouring pixels): the wave walks one
// shared tree path with a wave-uniform stack; a child is entered when any
// active lane's ray enters it within that lane's [tmin, closest t so far],
// and every lane keeps its own closest hit. Each lane's result is therefore
// the closest (t, instance, geometry, primitive) over a superset of the
// leaves its own traversal would reach, i.e. the same hit, computed with the
// same node, instance and triangle arithmetic as Tracer. Children are ordered
// by the entry distance of the first lane that enters them.
// ---------------------------------------------------------------------------
constexpr uint32_t kPacketStack = 128;  // wave-uniform stack entries (as deep as a lane stack)

template <int L>
__device__ inline void trace_packet(const SceneGPU& S, bool active, f3 wo, f3 wd, float tmin, float tmax, uint32_t rflags,
                                    HitRec& h, HH_LDS uint32_t* stack) {
    constexpr int W = layout_width(L);
    const HH_GLOBAL void* nodes = (const HH_GLOBAL void*)S.nodes;
    const HH_GLOBAL float4* tris = (const HH_GLOBAL float4*)S.tris;
    const HH_GLOBAL InstTrav* insts = (const HH_GLOBAL InstTrav*)S.inst_trav;
    const f3 winv = box_inv(wd);
    f3 o = wo, inv = winv;
    WT wt = wt_pre(wd);
    const bool alpha_on = (rflags & 1u) == 0u;
    h.found = false;
    h.t = tmax;
    h.u = h.v = 0.0f;
    h.inst = 0xFFFFFFFFu;
    h.tri = 0xFFFFFFFFu;
    uint32_t hgeom = 0xFFFFFFFFu, hprim = 0xFFFFFFFFu;
    const bool lane0 = __lane_id() == 0;
    uint32_t cur = S.tlas_root, sp = 0, trips = 0;
    int inst = -1;
    auto push = [&](uint32_t r) {
        if (sp < kPacketStack) {
            if (lane0) stack[sp] = r;
            sp++;
        } else {
            trips = kStackFull;  // ends the packet; counted below
        }
    };
    auto pop = [&]() { cur = sp == 0 ? kDone : (uint32_t)__builtin_amdgcn_readfirstlane((int)stack[--sp]); };
    while (cur != kDone && trips < kMaxTrips) {
        trips++;
        if ((cur & kLeafBit) == 0u) {
            float tn[W];
            uint32_t ref[W];
            node_test<L>(nodes, cur, o, inv, tmin, h.t, tn, ref);
            // uniform sort keys: entry distance of the first entering lane
            // (non-negative, so its bits order as unsigned), child index in
            // the low bits; 0xFFFFFFFF = no lane enters
            uint32_t key[W];
#pragma unroll
            for (int k = 0; k < W; k++) {
                const unsigned long long m = __ballot(active && tn[k] != INFINITY);
                if (m == 0ull) {
                    key[k] = 0xFFFFFFFFu;
                } else {
                    const int first = __ffsll((long long)m) - 1;
                    const uint32_t bits = (uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(tn[k]), first);
                    key[k] = (bits & ~7u) | (uint32_t)k;
                }
            }
#pragma unroll
            for (int i = 1; i < W; i++)
#pragma unroll
                for (int j = W - 1; j >= i; j--)
                    if (key[j] < key[j - 1]) {
                        const uint32_t t = key[j];
                        key[j] = key[j - 1];
                        key[j - 1] = t;
                    }
            if (key[0] == 0xFFFFFFFFu) {
                pop();
                continue;
            }
            // child reference by (uniform) index, as a select chain
            auto pick = [&](uint32_t i) {
                uint32_t r = ref[0];
#pragma unroll
                for (int k = 1; k < W; k++) r = i == (uint32_t)k ? ref[k] : r;
                return (uint32_t)__builtin_amdgcn_readfirstlane((int)r);
            };
#pragma unroll
            for (int k = W - 1; k >= 1; k--)
                if (key[k] != 0xFFFFFFFFu) push(pick(key[k] & 7u));
            cur = pick(key[0] & 7u);
        } else if (cur == kMarker) {
            o = wo;
            inv = winv;
            wt = wt_pre(wd);
            inst = -1;
            pop();
        } else if (inst < 0) {
            inst = (int)leaf_start(cur);
            const HH_GLOBAL InstTrav& it = insts[inst];
            f3 od;
            xform_ray(it.w2o, wo, wd, o, od);
            inv = box_inv(od);
            wt = wt_pre(od);
            push(kMarker);
            cur = (uint32_t)__builtin_amdgcn_readfirstlane((int)it.blas_root);
        } else {
            const uint32_t first = leaf_start(cur), ntri = leaf_count(cur);
            if (active) {
                for (uint32_t k = 0; k < ntri; k++) {
                    const uint32_t ti = first + k;
                    const float4 a = tris[3 * ti], b = tris[3 * ti + 1], c = tris[3 * ti + 2];
                    float t, u, v;
                    if (!tri_hit(wt, o, ld3(a), ld3(b), ld3(c), tmin, h.t, t, u, v)) continue;
                    const uint32_t prim = __float_as_uint(a.w);
                    const uint32_t gw = __float_as_uint(b.w);
                    const uint32_t geom = gw & 0x7FFFFFFFu;
                    if (h.found && t == h.t && !key_less((uint32_t)inst, geom, prim, h.inst, hgeom, hprim)) continue;
                    if (alpha_on && (gw & 0x80000000u) == 0u) {
                        if (hit_alpha(S, (uint32_t)inst, geom, ti, u, v) < 0.1f) continue;
                    }
                    h.found = true;
                    h.t = t;
                    h.u = u;
                    h.v = v;
                    h.inst = (uint32_t)inst;
                    h.tri = ti;
                    hgeom = geom;
                    hprim = prim;
                }
            }
            pop();
        }
    }
    if (cur != kDone && lane0) atomicAdd(&S.faults[trips == kStackFull ? kFaultPacketStack : kFaultTrips], 1ull);
}

}  // namespace hh
