// Karras binary radix tree over the sorted leaf codes and the bottom-up moment climb.
//
// Reference: bh_gravity.wgsl:110-184 (Karras build, duplicate codes tie-broken by leaf index) and
// :186-241 (single-pass climb). The reference engine replaced its climb with 32+ level-synchronous sweeps
// because WGSL cannot order the sibling reads (bh_gravity.wgsl:192-199, gpu_gravity.rs:484-519). HIP can:
// each climber publishes its node with a device-scope release fence before the arrival atomic, and the
// second arriver acquires before reading the sibling, so one pass suffices. The merge is symmetric in the
// two children, so the result does not depend on which child arrives second (deterministic).
#include "state.h"

namespace isph {

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
                                                uint32_t* __restrict__ parent, uint32_t* __restrict__ ready) {
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
    if (fc == lc) {
        split = (first + last) >> 1;
    } else {
        int cpfx = __clz(fc ^ lc);
        int step = last - first;
        do {
            step = (step + 1) >> 1;
            int ns = split + step;
            if (ns < last && __clz(fc ^ codes[ns]) > cpfx) split = ns;
        } while (step > 1);
    }
    uint32_t li = (split == first) ? (L - 1) + (uint32_t)split : (uint32_t)split;
    uint32_t ri = (split + 1 == last) ? (L - 1) + (uint32_t)(split + 1) : (uint32_t)(split + 1);
    left[ii] = li;
    right[ii] = ri;
    parent[li] = ii;
    parent[ri] = ii;
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

// One thread per leaf: bucket moments (bh_gravity.wgsl:200-223), then climb (:224-240).
__global__ __launch_bounds__(256) void k_upsweep(const float4* __restrict__ spos, const float* __restrict__ sh,
                                                 uint32_t n, uint32_t K, uint32_t L, const uint32_t* __restrict__ left,
                                                 const uint32_t* __restrict__ right,
                                                 const uint32_t* __restrict__ parent, uint32_t* ready, float4* nA,
                                                 float4* nB, float4* nC, float4* nQ0, float2* nQ1, float inv_theta,
                                                 int quad) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
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
    float4 q0 = make_float4(0, 0, 0, 0);
    float2 q1 = make_float2(0, 0);
    if (quad && end - start > 1) {
        for (uint32_t s = start; s < end; ++s) {
            float4 b = spos[s];
            add_shifted_quad(q0, q1, make_float4(0, 0, 0, 0), make_float2(0, 0), b.w, b.x - com.x, b.y - com.y,
                             b.z - com.z);
        }
    }
    nA[li] = make_float4(com.x, com.y, com.z, m);
    nB[li] = make_float4(lo.x, lo.y, lo.z, hmax);
    nC[li] = make_float4(hi.x, hi.y, hi.z, mac_radius(lo, hi, com, inv_theta));
    if (quad) {
        nQ0[li] = q0;
        nQ1[li] = q1;
    }
    __threadfence();  // release: this node's moments are visible device-wide before the arrival count

    uint32_t node = parent[li];
    while (node != kNoParent) {
        uint32_t arrived = atomicAdd(&ready[node], 1u);
        if (arrived == 0) break;  // first child: the sibling's climber finishes this node
        __threadfence();          // acquire: the sibling's published moments are read below
        uint32_t l = left[node], r = right[node];
        float4 al = nA[l], ar = nA[r];
        float4 bl = nB[l], br = nB[r];
        float4 cl = nC[l], cr = nC[r];
        float mm = al.w + ar.w;
        float3 cm = mm > 0.0f ? make_float3((al.x * al.w + ar.x * ar.w) / mm, (al.y * al.w + ar.y * ar.w) / mm,
                                            (al.z * al.w + ar.z * ar.w) / mm)
                              : make_float3(0.0f, 0.0f, 0.0f);
        float3 nlo = make_float3(fminf(bl.x, br.x), fminf(bl.y, br.y), fminf(bl.z, br.z));
        float3 nhi = make_float3(fmaxf(cl.x, cr.x), fmaxf(cl.y, cr.y), fmaxf(cl.z, cr.z));
        nA[node] = make_float4(cm.x, cm.y, cm.z, mm);
        nB[node] = make_float4(nlo.x, nlo.y, nlo.z, fmaxf(bl.w, br.w));
        nC[node] = make_float4(nhi.x, nhi.y, nhi.z, mac_radius(nlo, nhi, cm, inv_theta));
        if (quad) {
            float4 qa = make_float4(0, 0, 0, 0);
            float2 qb = make_float2(0, 0);
            add_shifted_quad(qa, qb, nQ0[l], nQ1[l], al.w,
                             al.x - cm.x, al.y - cm.y, al.z - cm.z);
            add_shifted_quad(qa, qb, nQ0[r], nQ1[r], ar.w,
                             ar.x - cm.x, ar.y - cm.y, ar.z - cm.z);
            nQ0[node] = qa;
            nQ1[node] = qb;
        }
        __threadfence();
        node = parent[node];
    }
}

void launch_karras(const Dev& d, hipStream_t s) {
    uint32_t threads = d.L > 1 ? d.L - 1 : 1;
    hipLaunchKernelGGL(k_karras, dim3((threads + 255) / 256), dim3(256), 0, s, d.lcode, d.L, d.left, d.right, d.parent,
                       d.ready);
}
void launch_upsweep(const Dev& d, const StepParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_upsweep, dim3((d.L + 255) / 256), dim3(256), 0, s, d.spos, d.sh, d.n, d.K, d.L, d.left,
                       d.right, d.parent, d.ready, d.nA, d.nB, d.nC, d.nQ0, d.nQ1, 1.0f / p.theta, p.quad);
}

}  // namespace isph
