// Karras binary radix tree over the sorted leaf codes and the bottom-up moment climb.
//
// Reference: bh_gravity.wgsl:110-184 (Karras build, duplicate codes tie-broken by leaf index; the split of
// all-equal-code ranges follows Karras 2012 instead of the shader's midpoint, see k_karras) and
// :186-241 (single-pass climb). The reference engine replaced its climb with 32+ level-synchronous sweeps
// because WGSL cannot order the sibling reads (bh_gravity.wgsl:192-199, gpu_gravity.rs:484-519). Here the
// climb runs level by level over chunks of leaves, every arrival inside one workgroup's LDS: level 1 takes
// chunks of 512 leaves (one leaf per thread), level k >= 2 chunks of 512 * 8^(k-1) leaves and the internal
// nodes whose range lies inside such a chunk but crosses a level-(k-1) chunk boundary. A node's children
// are either of its own level (they arrive through LDS) or complete from a lower level (read from HBM once).
// Each node is computed from its two children in a fixed left/right order, so the result does not depend
// on arrival order.
#include "state.h"

namespace isph {

// Upsweep levels: an internal node's level is the smallest k whose chunk of kLv1 * 8^(k-1) leaves holds its
// whole range. The nodes of one level-k chunk (k >= 2) all contain one of its seven inner level-(k-1)
// boundaries; the nodes containing one boundary are nested (one root path, at most 97 nodes for a 64-bit
// key with the 32-bit index tie-break), so a chunk has at most 7 * 97 < kSlots of them.
// Each workgroup also hands the next level the nodes of its chunk that belong to a higher level: such a
// node's id lies in the chunk and its range leaves it, so it contains the chunk's left or right boundary,
// and there are at most 2 * 97 < kCarry of them.
constexpr uint32_t kLv1 = kUpLeaves;
constexpr uint32_t kLvShift = 3;  // chunk growth per level: x8
constexpr uint32_t kSlots = 768;  // LDS slots per level-k workgroup
constexpr uint32_t kCarry = kUpCarry;  // carried higher-level ids per workgroup
constexpr uint32_t kNoSlot = 0xffffffffu;

__device__ __forceinline__ int clz_key(uint32_t x) { return __clz(x); }
__device__ __forceinline__ int clz_key(uint64_t x) { return __clzll(x); }

// Common-prefix length of the index-extended keys (code, index) of leaves i and j; -1 outside [0, L).
template <class Key>
__device__ __forceinline__ int delta_codes(const Key* __restrict__ codes, int L, int i, int j) {
    if (j < 0 || j >= L) return -1;
    Key ci = codes[i], cj = codes[j];
    if (ci == cj) return (int)(8 * sizeof(Key)) + __clz((uint32_t)i ^ (uint32_t)j);
    return clz_key(ci ^ cj);
}

// One thread per internal node (bh_gravity.wgsl:134-184), with the node's upsweep level. Also sets the
// root's parent (cs_tree_reset, :122-131, restricted to what is read later). Keys are the 30-bit leaf codes, or
// (instance << 30 | code) for an ensemble, where each instance's leaves then form one subtree.
template <class Key>
__global__ __launch_bounds__(256) void k_karras(const Key* __restrict__ codes, uint32_t L,
                                                uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                                                uint32_t* __restrict__ parent, uint2* __restrict__ nrange,
                                                SphNode* __restrict__ nS, uint8_t* __restrict__ nlevel) {
    uint32_t ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii == 0) parent[0] = kNoParent;
    if (ii + 1 >= L) return;  // L-1 internal nodes
    int i = (int)ii, Li = (int)L;
    int d = delta_codes(codes, Li, i, i + 1) > delta_codes(codes, Li, i, i - 1) ? 1 : -1;
    int dmin = delta_codes(codes, Li, i, i - d);
    int lmax = 2;
    while (delta_codes(codes, Li, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta_codes(codes, Li, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int first = min(i, j), last = max(i, j);
    int split = first;
    Key fc = codes[first], lc = codes[last];
    // findSplit on the index-extended key (Karras 2012). For first/last codes that differ this is exactly
    // the shader's search (:162-172); for an all-equal range the shader's midpoint split (:159-160)
    // disagrees with the index tie-break of determineRange and breaks the tree on runs of >= 3 equal
    // codes, so the extended delta is used there too.
    {
        int cpfx = (fc == lc) ? delta_codes(codes, Li, first, last) : clz_key(fc ^ lc);
        int step = last - first;
        do {
            step = (step + 1) >> 1;
            int ns = split + step;
            if (ns < last && delta_codes(codes, Li, first, ns) > cpfx) split = ns;
        } while (step > 1);
    }
    uint32_t li = (split == first) ? (L - 1) + (uint32_t)split : (uint32_t)split;
    uint32_t ri = (split + 1 == last) ? (L - 1) + (uint32_t)(split + 1) : (uint32_t)(split + 1);
    left[ii] = li;
    right[ii] = ri;
    parent[li] = ii;
    parent[ri] = ii;
    nrange[ii] = make_uint2((uint32_t)first, (uint32_t)last);
    nS[ii].lr = make_uint4(li, ri, (uint32_t)first, (uint32_t)last);
    uint32_t level = 1;
    for (uint64_t span = kLv1; (uint64_t)first / span != (uint64_t)last / span; span <<= kLvShift) ++level;
    nlevel[ii] = (uint8_t)level;
}

__device__ __forceinline__ float mac_radius(float3 lo, float3 hi, float3 com, float inv_theta) {
    // bh_gravity.wgsl:375-382: accept when dist >= |bmax - bmin| / theta + |centre - com|
    float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    float size = sqrtf(ex * ex + ey * ey + ez * ez);
    float cx = 0.5f * (lo.x + hi.x) - com.x, cy = 0.5f * (lo.y + hi.y) - com.y, cz = 0.5f * (lo.z + hi.z) - com.z;
    float delta = sqrtf(cx * cx + cy * cy + cz * cz);
    return size * inv_theta + delta;
}

// Quadrupoles are stored per unit mass, Qhat = Q / M with Q_ij = sum m (3 d_i d_j - |d|^2 delta_ij) about the
// node COM: in SI units Q itself (~M R^2 ~ 1e38 for a planet) would overflow f32, while Qhat ~ R^2 does not.
// Adds f * (Qhat_child + 3 d d - |d|^2 I) for a child (or particle, Qhat = 0) of mass fraction f at offset
// d from the parent COM.
__device__ __forceinline__ void add_shifted_quad(float4& q0, float2& q1, float4 cq0, float2 cq1, float f, float dx,
                                                 float dy, float dz) {
    float r2 = dx * dx + dy * dy + dz * dz;
    q0.x += f * (cq0.x + 3.0f * dx * dx - r2);
    q0.y += f * (cq0.y + 3.0f * dy * dy - r2);
    q0.z += f * (cq0.z + 3.0f * dz * dz - r2);
    q0.w += f * (cq0.w + 3.0f * dx * dy);
    q1.x += f * (cq1.x + 3.0f * dx * dz);
    q1.y += f * (cq1.y + 3.0f * dy * dz);
}

struct NodeMom {
    float4 a;   // com.xyz, mass
    float4 b;   // bmin.xyz, max h
    float4 c;   // bmax.xyz, MAC radius
    float4 q0;  // quadrupole per unit mass: xx, yy, zz, xy
    float2 q1;  // quadrupole per unit mass: xz, yz
};

struct TreeView {
    const uint32_t* left;
    const uint32_t* right;
    const uint32_t* parent;
    const uint2* nrange;
    uint4* nG;
    float4* nA;
    SphNode* nS;
    float4* nQ0;
    float2* nQ1;
};

__device__ __forceinline__ void store_node(const TreeView& tv, uint32_t i, const NodeMom& m, int quad, uint32_t l,
                                           uint32_t r) {
    tv.nA[i] = m.a;
    tv.nS[i].lo = m.b;
    tv.nS[i].hi = m.c;
    tv.nG[2 * i] = make_uint4(__float_as_uint(m.a.x), __float_as_uint(m.a.y), __float_as_uint(m.a.z),
                              __float_as_uint(m.a.w));
    tv.nG[2 * i + 1] = make_uint4(__float_as_uint(m.c.w), l, r, 0u);
    if (quad) {
        tv.nQ0[i] = m.q0;
        tv.nQ1[i] = m.q1;
    }
}

__device__ __forceinline__ NodeMom load_mom(const TreeView& tv, uint32_t i, bool quad) {
    NodeMom m;
    m.a = tv.nA[i];
    m.b = tv.nS[i].lo;
    m.c = tv.nS[i].hi;
    m.q0 = quad ? tv.nQ0[i] : make_float4(0, 0, 0, 0);
    m.q1 = quad ? tv.nQ1[i] : make_float2(0, 0);
    return m;
}

// Parent moments from the left and right child (bh_gravity.wgsl:229-238: mass-weighted COM, unioned AABB;
// plus the max smoothing length, the MAC radius and the shifted quadrupoles).
__device__ __forceinline__ NodeMom merge_moms(const NodeMom& L_, const NodeMom& R_, float inv_theta, bool quad) {
    const float4 al = L_.a, ar = R_.a, bl = L_.b, br = R_.b, cl = L_.c, cr = R_.c;
    NodeMom m;
    float mm = al.w + ar.w;
    float3 cm = mm > 0.0f ? make_float3((al.x * al.w + ar.x * ar.w) / mm, (al.y * al.w + ar.y * ar.w) / mm,
                                        (al.z * al.w + ar.z * ar.w) / mm)
                          : make_float3(0.0f, 0.0f, 0.0f);
    float3 nlo = make_float3(fminf(bl.x, br.x), fminf(bl.y, br.y), fminf(bl.z, br.z));
    float3 nhi = make_float3(fmaxf(cl.x, cr.x), fmaxf(cl.y, cr.y), fmaxf(cl.z, cr.z));
    m.a = make_float4(cm.x, cm.y, cm.z, mm);
    m.b = make_float4(nlo.x, nlo.y, nlo.z, fmaxf(bl.w, br.w));
    m.c = make_float4(nhi.x, nhi.y, nhi.z, mac_radius(nlo, nhi, cm, inv_theta));
    m.q0 = make_float4(0, 0, 0, 0);
    m.q1 = make_float2(0, 0);
    if (quad && mm > 0.0f) {
        const float fl = al.w / mm, fr = ar.w / mm;
        add_shifted_quad(m.q0, m.q1, L_.q0, L_.q1, fl, al.x - cm.x, al.y - cm.y, al.z - cm.z);
        add_shifted_quad(m.q0, m.q1, R_.q0, R_.q1, fr, ar.x - cm.x, ar.y - cm.y, ar.z - cm.z);
    }
    return m;
}

// LDS state of one workgroup's climb: per slot (an internal node of this level) the moments of its two
// children as they arrive, an arrival counter, the node, its children and the slot of its parent
// (kNoSlot when the parent belongs to a higher level; bit 31 = right child).
template <bool QUAD>
struct Climb {
    float4 *A, *B, *C, *Q0;
    float2* Q1;
    uint32_t *cnt, *node, *par;
    uint2* child;

    __device__ __forceinline__ void put(uint32_t s, uint32_t side, const NodeMom& m) const {
        const uint32_t e = 2 * s + side;
        A[e] = m.a;
        B[e] = m.b;
        C[e] = m.c;
        if (QUAD) {
            Q0[e] = m.q0;
            Q1[e] = m.q1;
        }
    }
    __device__ __forceinline__ NodeMom get(uint32_t e) const {
        NodeMom m;
        m.a = A[e];
        m.b = B[e];
        m.c = C[e];
        m.q0 = QUAD ? Q0[e] : make_float4(0, 0, 0, 0);
        m.q1 = QUAD ? Q1[e] : make_float2(0, 0);
        return m;
    }
    // Deliver child moments `m` to slot s; the second arrival merges, writes the node to HBM and carries
    // on to the parent slot. Arrivals are LDS atomics with workgroup-scope acquire/release.
    __device__ __forceinline__ void deliver(uint32_t s, uint32_t side, NodeMom m, const TreeView& tv,
                                           float inv_theta) const {
        while (true) {
            put(s, side, m);
            if (__hip_atomic_fetch_add(&cnt[s], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u) return;
            m = merge_moms(get(2 * s), get(2 * s + 1), inv_theta, QUAD);
            const uint2 ch = child[s];
            store_node(tv, node[s], m, QUAD, ch.x, ch.y);
            const uint32_t p = par[s];
            if (p == kNoSlot) return;
            s = p & 0x7fffffffu;
            side = p >> 31;
        }
    }
};

#define ISPH_CLIMB_LDS(QUAD, N)                                              \
    __shared__ float4 sA[2 * (N)], sB[2 * (N)], sC[2 * (N)];                 \
    __shared__ float4 sQ0[QUAD ? 2 * (N) : 1];                               \
    __shared__ float2 sQ1[QUAD ? 2 * (N) : 1];                               \
    __shared__ uint32_t sCnt[N], sNode[N], sPar[N];                          \
    __shared__ uint2 sChild[N];                                              \
    const Climb<QUAD> cl{sA, sB, sC, sQ0, sQ1, sCnt, sNode, sPar, sChild}

// Level 1: one thread per leaf of a kLv1-leaf chunk (bucket moments, bh_gravity.wgsl:200-223), then the
// climb (:224-240) through the chunk's level-1 nodes, whose ids are the chunk's leaf indices (a Karras
// node's id is one end of its range), so slot = id - first leaf of the chunk.
template <bool QUAD>
__global__ __launch_bounds__(kLv1) void k_upsweep_lv1(const float4* __restrict__ spos, const float* __restrict__ sh,
                                                      uint32_t n, uint32_t K, uint32_t L, TreeView tv,
                                                      const SphNode* __restrict__ nS,
                                                      const uint8_t* __restrict__ nlevel, float inv_theta,
                                                      uint32_t* __restrict__ carry, uint32_t* __restrict__ ncarry) {
    ISPH_CLIMB_LDS(QUAD, kLv1);
    __shared__ uint32_t nup;
    const uint32_t c0 = blockIdx.x * kLv1, t = threadIdx.x;
    if (t == 0) nup = 0u;
    __syncthreads();
    auto slot_of_parent = [&](uint32_t child) {
        const uint32_t p = tv.parent[child];
        if (p == kNoParent || nlevel[p] != 1) return kNoSlot;
        return (p - c0) | (nS[p].lr.x == child ? 0u : 0x80000000u);
    };
    // slot metadata of internal node c0 + t
    const uint32_t id = c0 + t;
    sCnt[t] = 0u;
    if (id + 1 < L) {
        if (nlevel[id] == 1) {
            sNode[t] = id;
            const uint4 c4 = nS[id].lr;
            sChild[t] = make_uint2(c4.x, c4.y);
            sPar[t] = slot_of_parent(id);
        } else {
            const uint32_t e = atomicAdd(&nup, 1u);
            if (e < kCarry) carry[(size_t)blockIdx.x * kCarry + e] = id;
        }
    }
    // leaf c0 + t
    const uint32_t c = c0 + t;
    uint32_t target = kNoSlot;
    NodeMom lm;
    if (c < L) {
        const uint32_t start = c * K, end = min(start + K, n);
        float m = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f, hmax = 0.0f;
        float3 lo = make_float3(1.0e30f, 1.0e30f, 1.0e30f), hi = make_float3(-1.0e30f, -1.0e30f, -1.0e30f);
        for (uint32_t s = start; s < end; ++s) {
            float4 b = spos[s];
            m += b.w;
            mx += b.x * b.w;
            my += b.y * b.w;
            mz += b.z * b.w;
            lo = make_float3(fminf(lo.x, b.x), fminf(lo.y, b.y), fminf(lo.z, b.z));
            hi = make_float3(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z));
            if (sh) hmax = fmaxf(hmax, sh[s]);  // gravity-only builds carry no smoothing lengths
        }
        float3 com;
        if (end - start == 1) com = make_float3(spos[start].x, spos[start].y, spos[start].z);
        else com = m > 0.0f ? make_float3(mx / m, my / m, mz / m) : make_float3(0.0f, 0.0f, 0.0f);
        lm.q0 = make_float4(0, 0, 0, 0);
        lm.q1 = make_float2(0, 0);
        if (QUAD && end - start > 1 && m > 0.0f) {
            for (uint32_t s = start; s < end; ++s) {
                float4 b = spos[s];
                add_shifted_quad(lm.q0, lm.q1, make_float4(0, 0, 0, 0), make_float2(0, 0), b.w / m, b.x - com.x,
                                 b.y - com.y, b.z - com.z);
            }
        }
        lm.a = make_float4(com.x, com.y, com.z, m);
        lm.b = make_float4(lo.x, lo.y, lo.z, hmax);
        lm.c = make_float4(hi.x, hi.y, hi.z, mac_radius(lo, hi, com, inv_theta));
        const uint32_t li = (L - 1) + c;
        store_node(tv, li, lm, QUAD, 0u, 0u);
        target = slot_of_parent(li);
    }
    __syncthreads();
    if (t == 0) ncarry[blockIdx.x] = min(nup, kCarry);
    if (target != kNoSlot) cl.deliver(target & 0x7fffffffu, target >> 31, lm, tv, inv_theta);
}

// Level k >= 2: the workgroup of a level-k chunk takes its level-k nodes from the ids that its eight
// level-(k-1) chunks carried up, gives each an LDS slot (`cslot` maps node -> slot for the parent links),
// carries the rest on, reads the children that a lower level completed from HBM and climbs.
template <bool QUAD>
__global__ __launch_bounds__(kSlots) void k_upsweep_lvk(uint32_t L, uint32_t level, uint32_t nbelow, TreeView tv,
                                                        const SphNode* __restrict__ nS,
                                                        const uint8_t* __restrict__ nlevel,
                                                        uint32_t* __restrict__ cslot,
                                                        const uint32_t* __restrict__ carry_in,
                                                        const uint32_t* __restrict__ ncarry_in,
                                                        uint32_t* __restrict__ carry_out,
                                                        uint32_t* __restrict__ ncarry_out, float inv_theta) {
    ISPH_CLIMB_LDS(QUAD, kSlots);
    __shared__ uint32_t nslots, nup;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        nslots = 0u;
        nup = 0u;
    }
    __syncthreads();
    for (uint32_t x = t; x < (kCarry << kLvShift); x += blockDim.x) {
        const uint32_t below = (blockIdx.x << kLvShift) + x / kCarry, e = x % kCarry;
        if (below >= nbelow || e >= ncarry_in[below]) continue;
        const uint32_t id = carry_in[(size_t)below * kCarry + e];
        if (nlevel[id] == level) {
            const uint32_t s = atomicAdd(&nslots, 1u);
            if (s < kSlots) {
                sNode[s] = id;
                cslot[id] = s;
            }
        } else {
            const uint32_t o = atomicAdd(&nup, 1u);
            if (o < kCarry) carry_out[(size_t)blockIdx.x * kCarry + o] = id;
        }
    }
    __syncthreads();
    if (t == 0) ncarry_out[blockIdx.x] = min(nup, kCarry);
    const uint32_t ns = min(nslots, kSlots);
    for (uint32_t s = t; s < ns; s += blockDim.x) {
        const uint32_t id = sNode[s];
        sCnt[s] = 0u;
        const uint4 c4 = nS[id].lr;
        sChild[s] = make_uint2(c4.x, c4.y);
        const uint32_t p = tv.parent[id];
        uint32_t ps = kNoSlot;
        if (p != kNoParent && nlevel[p] == level) {
            const uint32_t cs = cslot[p];
            if (cs < ns) ps = cs | (nS[p].lr.x == id ? 0u : 0x80000000u);
        }
        sPar[s] = ps;
    }
    __syncthreads();
    for (uint32_t s = t; s < ns; s += blockDim.x) {
        const uint2 ch = sChild[s];
#pragma unroll
        for (uint32_t side = 0; side < 2; ++side) {
            const uint32_t c = side ? ch.y : ch.x;
            if (c + 1 < L && nlevel[c] == level) continue;  // arrives by its own climb
            cl.deliver(s, side, load_mom(tv, c, QUAD), tv, inv_theta);
        }
    }
}

// Walk groups: the highest nodes holding <= `maxc` particles whose box diagonal is at most `extent` times
// the largest smoothing length of the scene (a leaf always qualifies). Both conditions only get easier down
// the tree, so these nodes partition the leaves into spatially compact, Morton-contiguous particle runs
// (runs of consecutive indices can straddle a Morton seam and span distant regions).
// Instance context of the group construction: the length unit of each node's limit comes from its
// instance's root (the whole tree for one system), and no group may span two instances.
struct GroupCtx {
    const SphNode* nS;
    const uint32_t* sinst;  // sorted instance ids, or nullptr for one system
    const uint32_t* iroot;
    const uint32_t* ioff;
    uint32_t n;
    float extent;
    // Largest h of the instance (carried by its root box), or for gravity-only builds without smoothing
    // lengths the root diagonal / cbrt(n) (about twice the mean spacing).
    __device__ __forceinline__ float lim2(uint32_t first_particle) const {
        uint32_t root = 0, cnt = n;
        if (sinst) {
            const uint32_t k = sinst[first_particle];
            root = iroot[k];
            cnt = ioff[k + 1] - ioff[k];
        }
        const float4 b0 = nS[root].lo, c0 = nS[root].hi;
        const float ex0 = c0.x - b0.x, ey0 = c0.y - b0.y, ez0 = c0.z - b0.z;
        const float spacing = sqrtf(ex0 * ex0 + ey0 * ey0 + ez0 * ez0) / cbrtf((float)cnt);
        const float lim = extent * (b0.w > 0.0f ? b0.w : spacing);
        return lim * lim;
    }
};

__device__ __forceinline__ bool group_ok(uint32_t v, uint32_t K, uint32_t L, const uint2* nrange, const GroupCtx& g,
                                         uint32_t maxc, uint32_t& start, uint32_t& count) {
    uint32_t first, last;
    bool leaf = v >= L - 1;
    if (leaf) {
        first = last = v - (L - 1);
    } else {
        uint2 r = nrange[v];
        first = r.x;
        last = r.y;
    }
    start = first * K;
    uint32_t end = min((last + 1) * K, g.n);
    count = end - start;
    if (leaf) return true;
    if (count > maxc) return false;
    if (g.sinst && g.sinst[start] != g.sinst[end - 1]) return false;
    float4 b = g.nS[v].lo, c = g.nS[v].hi;
    float ex = c.x - b.x, ey = c.y - b.y, ez = c.z - b.z;
    return ex * ex + ey * ey + ez * ez <= g.lim2(start);
}

// Mark each group at its first particle (gmark[start] = count); gmark is all-zero on entry.
__global__ __launch_bounds__(256) void k_group_mark(uint32_t K, uint32_t L, const uint32_t* __restrict__ parent,
                                                    const uint2* __restrict__ nrange, GroupCtx g, uint32_t maxc,
                                                    uint32_t* __restrict__ gmark) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * L - 1) return;
    uint32_t start, count, ps, pc;
    if (!group_ok(v, K, L, nrange, g, maxc, start, count)) return;
    uint32_t p = parent[v];
    if (p != kNoParent && group_ok(p, K, L, nrange, g, maxc, ps, pc)) return;
    gmark[start] = count;
}

// Compaction of the marks in particle (Morton) order: per 64-particle block count, exclusive scan of the
// counts, emit. The emit pass clears the marks for the next build and stores the group total.
__global__ __launch_bounds__(256) void k_group_count(uint32_t n, const uint32_t* __restrict__ gmark,
                                                     uint32_t* __restrict__ bcnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t b = __ballot(i < n && gmark[i] != 0u);
    if (lane_id() == 0 && i / kWave < (n + kWave - 1) / kWave) bcnt[i / kWave] = (uint32_t)__popcll(b);
}
__global__ __launch_bounds__(256) void k_group_emit(uint32_t n, uint32_t* __restrict__ gmark,
                                                    const uint32_t* __restrict__ boff, uint2* __restrict__ groups,
                                                    uint32_t* __restrict__ total) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = i < n ? gmark[i] : 0u;
    uint64_t b = __ballot(c != 0u);
    uint32_t nb = (n + kWave - 1) / kWave;
    if (i / kWave >= nb) return;  // whole wave
    uint32_t off = boff[i / kWave];
    if (c) {
        groups[off + prefix_count(b)] = make_uint2(i, c);
        gmark[i] = 0u;
    }
    if (i / kWave == nb - 1 && lane_id() == 0) *total = off + (uint32_t)__popcll(b);
}

void launch_groups(const Dev& d, const StepParams& p, uint32_t maxc, uint2* out, uint32_t* total, hipStream_t s) {
    uint32_t nodes = 2 * d.L - 1;
    uint32_t nb = (d.n + kWave - 1) / kWave;
    GroupCtx g{d.nS, d.sinst(), d.iroot, d.ioff, d.n, p.group_extent};
    hipLaunchKernelGGL(k_group_mark, dim3((nodes + 255) / 256), dim3(256), 0, s, d.K, d.L, d.parent, d.nrange, g,
                       maxc, d.gflag);
    hipLaunchKernelGGL(k_group_count, dim3((d.n + 255) / 256), dim3(256), 0, s, d.n, d.gflag, d.gscan);
    launch_exclusive_scan(d.gscan, nb, s);
    hipLaunchKernelGGL(k_group_emit, dim3((d.n + 255) / 256), dim3(256), 0, s, d.n, d.gflag, d.gscan, out, total);
}

void launch_karras(const Dev& d, hipStream_t s) {
    uint32_t threads = d.L > 1 ? d.L - 1 : 1;
    if (d.ninst > 1)
        hipLaunchKernelGGL(k_karras<uint64_t>, dim3((threads + 255) / 256), dim3(256), 0, s, d.lkey, d.L, d.left,
                           d.right, d.parent, d.nrange, d.nS, d.nlevel);
    else
        hipLaunchKernelGGL(k_karras<uint32_t>, dim3((threads + 255) / 256), dim3(256), 0, s, d.lcode, d.L, d.left,
                           d.right, d.parent, d.nrange, d.nS, d.nlevel);
}
void launch_upsweep(const Dev& d, const StepParams& p, hipStream_t s) {
    TreeView tv{d.left, d.right, d.parent, d.nrange, d.nG, d.nA, d.nS, d.nQ0, d.nQ1};
    const float inv_theta = 1.0f / p.theta;
    // carried ids: level-1 workgroups write list 0, level k reads list (k % 2) ^ 1 ... alternating halves
    uint32_t* lists[2] = {d.carry, d.carry + (size_t)d.carry_blocks * kCarry};
    uint32_t* counts[2] = {d.ncarry, d.ncarry + d.carry_blocks};
    auto go = [&](auto quad) {
        constexpr bool Q = decltype(quad)::value;
        uint32_t nbelow = (d.L + kLv1 - 1) / kLv1;
        hipLaunchKernelGGL(k_upsweep_lv1<Q>, dim3(nbelow), dim3(kLv1), 0, s, d.spos, d.sh, d.n, d.K, d.L, tv, d.nS,
                           d.nlevel, inv_theta, lists[0], counts[0]);
        uint32_t level = 2, in = 0;
        for (uint64_t span = (uint64_t)kLv1 << kLvShift; (span >> kLvShift) < d.L; span <<= kLvShift, ++level) {
            const uint32_t blocks = (uint32_t)((d.L + span - 1) / span);
            hipLaunchKernelGGL(k_upsweep_lvk<Q>, dim3(blocks), dim3(kSlots), 0, s, d.L, level, nbelow, tv, d.nS,
                               d.nlevel, d.cslot, lists[in], counts[in], lists[in ^ 1], counts[in ^ 1], inv_theta);
            nbelow = blocks;
            in ^= 1;
        }
    };
    if (p.quad)
        go(std::true_type{});
    else
        go(std::false_type{});
}

}  // namespace isph
