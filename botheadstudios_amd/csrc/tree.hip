// Karras binary radix tree over the sorted leaf codes and the bottom-up moment climb.
//
// Reference: bh_gravity.wgsl:110-184 (Karras build, duplicate codes tie-broken by leaf index; the split of
// all-equal-code ranges follows Karras 2012 instead of the shader's midpoint, see k_karras) and
// :186-241 (single-pass climb). The reference engine replaced its climb with 32+ level-synchronous sweeps
// because WGSL cannot order the sibling reads (bh_gravity.wgsl:192-199, gpu_gravity.rs:484-519). Here the
// climb runs in two launches whose arrivals are ordered with workgroup-scope acquire/release only (a
// device-scope fence writes back the XCD-local L2 on every climber): pass 1 climbs inside chunks of 256
// leaves, pass 2 (one workgroup) finishes the nodes that straddle chunks. Each node is computed from its
// two children in a fixed left/right order, so the result does not depend on arrival order.
#include "state.h"

namespace isph {

// Leaves per upsweep chunk: internal nodes whose leaf range lies inside one chunk are finished by that
// chunk's workgroup; the few whose range crosses a chunk boundary by one final workgroup.
constexpr uint32_t kUpChunk = 256;

__device__ __forceinline__ int delta_codes(const uint32_t* __restrict__ codes, int L, int i, int j) {
    if (j < 0 || j >= L) return -1;
    uint32_t ci = codes[i], cj = codes[j];
    if (ci == cj) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(ci ^ cj);
}

// One thread per internal node (bh_gravity.wgsl:134-184). Also clears the climb counters and the root's
// parent (cs_tree_reset, :122-131, restricted to what is read later).
__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ codes, uint32_t L,
                                                uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                                                uint32_t* __restrict__ parent, uint32_t* __restrict__ ready,
                                                uint2* __restrict__ nrange, uint32_t* __restrict__ cross,
                                                uint32_t* scal) {
    uint32_t ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii == 0) parent[0] = kNoParent;
    if (ii + 1 >= L) return;  // L-1 internal nodes
    int i = (int)ii, Li = (int)L;
    ready[ii] = 0;
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
    uint32_t fc = codes[first], lc = codes[last];
    // findSplit on the index-extended key (Karras 2012). For first/last codes that differ this is exactly
    // the shader's search (:162-172); for an all-equal range the shader's midpoint split (:159-160)
    // disagrees with the index tie-break of determineRange and breaks the tree on runs of >= 3 equal
    // codes, so the extended delta is used there too.
    {
        int cpfx = (fc == lc) ? delta_codes(codes, Li, first, last) : __clz(fc ^ lc);
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
    if ((uint32_t)first / kUpChunk != (uint32_t)last / kUpChunk) cross[atomicAdd(&scal[7], 1u)] = ii;
}

__device__ __forceinline__ float mac_radius(float3 lo, float3 hi, float3 com, float inv_theta) {
    // bh_gravity.wgsl:375-382: accept when dist >= |bmax - bmin| / theta + |centre - com|
    float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    float size = sqrtf(ex * ex + ey * ey + ez * ez);
    float cx = 0.5f * (lo.x + hi.x) - com.x, cy = 0.5f * (lo.y + hi.y) - com.y, cz = 0.5f * (lo.z + hi.z) - com.z;
    float delta = sqrtf(cx * cx + cy * cy + cz * cz);
    return size * inv_theta + delta;
}

__device__ __forceinline__ void add_shifted_quad(float4& q0, float2& q1, float4 cq0, float2 cq1, float m, float dx,
                                                 float dy, float dz) {
    // Q_ij about the parent COM = Q_child + m (3 d_i d_j - |d|^2 delta_ij), d = child com - parent com.
    float r2 = dx * dx + dy * dy + dz * dz;
    q0.x += cq0.x + m * (3.0f * dx * dx - r2);
    q0.y += cq0.y + m * (3.0f * dy * dy - r2);
    q0.z += cq0.z + m * (3.0f * dz * dz - r2);
    q0.w += cq0.w + m * (3.0f * dx * dy);
    q1.x += cq1.x + m * (3.0f * dx * dz);
    q1.y += cq1.y + m * (3.0f * dy * dz);
}

struct NodeMom {
    float4 a;   // com.xyz, mass
    float4 b;   // bmin.xyz, max h
    float4 c;   // bmax.xyz, MAC radius
    float4 q0;  // quadrupole xx, yy, zz, xy
    float2 q1;  // quadrupole xz, yz
};

struct TreeView {
    const uint32_t* left;
    const uint32_t* right;
    const uint32_t* parent;
    const uint2* nrange;
    uint32_t* ready;
    float4* nA;
    float4* nB;
    float4* nC;
    float4* nQ0;
    float2* nQ1;
};

__device__ __forceinline__ void store_node(const TreeView& tv, uint32_t i, const NodeMom& m, int quad) {
    tv.nA[i] = m.a;
    tv.nB[i] = m.b;
    tv.nC[i] = m.c;
    if (quad) {
        tv.nQ0[i] = m.q0;
        tv.nQ1[i] = m.q1;
    }
}

// Merge the two children of `node` (bh_gravity.wgsl:229-238: mass-weighted COM, unioned AABB; plus the
// max smoothing length, the MAC radius and the shifted quadrupoles).
__device__ __forceinline__ NodeMom merge_children(const TreeView& tv, uint32_t node, float inv_theta, int quad) {
    uint32_t l = tv.left[node], r = tv.right[node];
    float4 al = tv.nA[l], ar = tv.nA[r];
    float4 bl = tv.nB[l], br = tv.nB[r];
    float4 cl = tv.nC[l], cr = tv.nC[r];
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
    if (quad) {
        add_shifted_quad(m.q0, m.q1, tv.nQ0[l], tv.nQ1[l], al.w, al.x - cm.x, al.y - cm.y, al.z - cm.z);
        add_shifted_quad(m.q0, m.q1, tv.nQ0[r], tv.nQ1[r], ar.w, ar.x - cm.x, ar.y - cm.y, ar.z - cm.z);
    }
    return m;
}

// Arrival at an internal node with workgroup-scope acquire/release: the second arriver merges and
// returns true. Every arrival at a given node comes from the same workgroup (the chunk's in pass 1, the
// single workgroup of pass 2), so no device-wide cache write-back is needed.
__device__ __forceinline__ bool second_arrival(uint32_t* ready, uint32_t node) {
    return __hip_atomic_fetch_add(&ready[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u;
}

__device__ __forceinline__ bool crosses(const TreeView& tv, uint32_t node) {
    uint2 r = tv.nrange[node];
    return r.x / kUpChunk != r.y / kUpChunk;
}

// Pass 1: one thread per leaf (bucket moments, bh_gravity.wgsl:200-223), climbing (:224-240) through the
// internal nodes whose range stays inside this workgroup's chunk of kUpChunk leaves.
__global__ __launch_bounds__(kUpChunk) void k_upsweep_chunk(const float4* __restrict__ spos,
                                                           const float* __restrict__ sh, uint32_t n, uint32_t K,
                                                           uint32_t L, TreeView tv, float inv_theta, int quad) {
    uint32_t c = blockIdx.x * kUpChunk + threadIdx.x;
    if (c >= L) return;
    uint32_t li = (L - 1) + c;
    uint32_t start = c * K, end = min(start + K, n);
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
        hmax = fmaxf(hmax, sh[s]);
    }
    float3 com;
    if (end - start == 1) com = make_float3(spos[start].x, spos[start].y, spos[start].z);
    else com = m > 0.0f ? make_float3(mx / m, my / m, mz / m) : make_float3(0.0f, 0.0f, 0.0f);
    NodeMom lm;
    lm.q0 = make_float4(0, 0, 0, 0);
    lm.q1 = make_float2(0, 0);
    if (quad && end - start > 1) {
        for (uint32_t s = start; s < end; ++s) {
            float4 b = spos[s];
            add_shifted_quad(lm.q0, lm.q1, make_float4(0, 0, 0, 0), make_float2(0, 0), b.w, b.x - com.x,
                             b.y - com.y, b.z - com.z);
        }
    }
    lm.a = make_float4(com.x, com.y, com.z, m);
    lm.b = make_float4(lo.x, lo.y, lo.z, hmax);
    lm.c = make_float4(hi.x, hi.y, hi.z, mac_radius(lo, hi, com, inv_theta));
    store_node(tv, li, lm, quad);

    uint32_t node = tv.parent[li];
    while (node != kNoParent && !crosses(tv, node)) {
        if (!second_arrival(tv.ready, node)) break;  // first child: the sibling's climber finishes it
        NodeMom nm = merge_children(tv, node, inv_theta, quad);
        store_node(tv, node, nm, quad);
        node = tv.parent[node];
    }
}

// Pass 2: one workgroup finishes the chunk-crossing internal nodes. Their non-crossing children are
// complete (pass 1, previous launch), so each crossing node receives one arrival per such child here and
// the climb continues to the root through crossing nodes only.
__global__ __launch_bounds__(1024) void k_upsweep_top(const uint32_t* __restrict__ cross, const uint32_t* scal,
                                                      uint32_t L, TreeView tv, float inv_theta, int quad) {
    const uint32_t count = scal[7];
    for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) {
        const uint32_t node = cross[k];
        const uint32_t ch[2] = {tv.left[node], tv.right[node]};
        for (int side = 0; side < 2; ++side) {
            const uint32_t child = ch[side];
            if (child < L - 1 && crosses(tv, child)) continue;  // a crossing child arrives by its own climb
            uint32_t at = node;
            while (at != kNoParent) {
                if (!second_arrival(tv.ready, at)) break;
                NodeMom nm = merge_children(tv, at, inv_theta, quad);
                store_node(tv, at, nm, quad);
                at = tv.parent[at];
            }
        }
    }
}

void launch_karras(const Dev& d, hipStream_t s) {
    uint32_t threads = d.L > 1 ? d.L - 1 : 1;
    hipLaunchKernelGGL(k_karras, dim3((threads + 255) / 256), dim3(256), 0, s, d.lcode, d.L, d.left, d.right, d.parent,
                       d.ready, d.nrange, d.cross, d.scal);
}
void launch_upsweep(const Dev& d, const StepParams& p, hipStream_t s) {
    TreeView tv{d.left, d.right, d.parent, d.nrange, d.ready, d.nA, d.nB, d.nC, d.nQ0, d.nQ1};
    float inv_theta = 1.0f / p.theta;
    hipLaunchKernelGGL(k_upsweep_chunk, dim3((d.L + kUpChunk - 1) / kUpChunk), dim3(kUpChunk), 0, s, d.spos, d.sh,
                       d.n, d.K, d.L, tv, inv_theta, p.quad);
    if (d.L > kUpChunk)
        hipLaunchKernelGGL(k_upsweep_top, dim3(1), dim3(1024), 0, s, d.cross, d.scal, d.L, tv, inv_theta, p.quad);
}

}  // namespace isph
