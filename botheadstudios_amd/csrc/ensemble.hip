// Ensembles: independent systems advanced by the same launches.
//
// Instance k owns the particles [ioff[k], ioff[k+1]) of the upload order. Each instance gets its own
// bounding box and Morton codes (bh_gravity.wgsl:49-108 per instance); the sort key is (instance, code),
// so after the sort the instances stay contiguous and Morton-ordered inside. The Karras tree over the
// 64-bit leaf keys (instance << 30 | code) has one subtree per instance (the keys sharing an instance
// prefix), whose root is found here; every walk starts at its group's instance root, so no interaction
// crosses instances. Each instance has its own Courant dt and simulated time.
#include "state.h"

namespace isph {

static inline uint32_t cdiv_e(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Per-instance bbox keys: waves whose particles all belong to one instance reduce first (the common case,
// instances are contiguous), others fall back to per-lane atomics.
__global__ __launch_bounds__(256) void k_bbox_seg(const float4* __restrict__ pos, uint32_t n,
                                                  const uint32_t* __restrict__ inst, uint32_t* ibox) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const uint32_t k = valid ? inst[i] : 0xffffffffu;
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    if (valid) {
        const float4 p = pos[i];
        lo[0] = hi[0] = bbox_key(p.x);
        lo[1] = hi[1] = bbox_key(p.y);
        lo[2] = hi[2] = bbox_key(p.z);
    }
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(k);
    const bool uniform = __ballot(valid && k != k0) == 0ull && __ballot(valid) != 0ull;
    if (uniform) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                lo[c] = min(lo[c], (uint32_t)__shfl_xor((int)lo[c], o));
                hi[c] = max(hi[c], (uint32_t)__shfl_xor((int)hi[c], o));
            }
        }
        if (lane_id() == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                atomicMin(&ibox[6 * k0 + c], lo[c]);
                atomicMax(&ibox[6 * k0 + 3 + c], hi[c]);
            }
        }
    } else if (valid) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            atomicMin(&ibox[6 * k + c], lo[c]);
            atomicMax(&ibox[6 * k + 3 + c], hi[c]);
        }
    }
}

__device__ __forceinline__ uint32_t expand_bits_e(uint32_t v) {  // bh_gravity.wgsl:82-89
    v &= 0x000003ffu;
    v = (v * 0x00010001u) & 0xff0000ffu;
    v = (v * 0x00000101u) & 0x0f00f00fu;
    v = (v * 0x00000011u) & 0xc30c30c3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Morton codes against the particle's instance box (the single-system k_morton, per instance).
__global__ __launch_bounds__(256) void k_morton_seg(const float4* __restrict__ pos, uint32_t n,
                                                    const uint32_t* __restrict__ inst,
                                                    const uint32_t* __restrict__ ibox, uint32_t* __restrict__ codes,
                                                    uint32_t* __restrict__ order, uint32_t* __restrict__ mcode) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* b = ibox + 6 * inst[i];
    float lo[3], ext[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = bbox_decode(b[k]);
        float hi = bbox_decode(b[3 + k]);
        ext[k] = fmaxf(__fsub_rn(hi, lo[k]), 1.0e-30f);
    }
    float4 p = pos[i];
    float c[3] = {p.x, p.y, p.z};
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float u = __fdiv_rn(__fsub_rn(c[k], lo[k]), ext[k]);
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        float f = floorf(__fmul_rn(u, 1024.0f));
        f = fminf(fmaxf(f, 0.0f), 1023.0f);
        q[k] = (uint32_t)f;
    }
    const uint32_t code = expand_bits_e(q[0]) * 4u + expand_bits_e(q[1]) * 2u + expand_bits_e(q[2]);
    codes[i] = code;
    mcode[i] = code;
    order[i] = i;
}

// Before the last (instance) radix pass: the instance of every element in the current order.
__global__ __launch_bounds__(256) void k_inst_key(uint32_t n, const uint32_t* __restrict__ vals,
                                                  const uint32_t* __restrict__ inst, uint32_t* __restrict__ kout,
                                                  uint32_t* __restrict__ vout) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t o = vals[t];
    kout[t] = inst[o];
    vout[t] = o;
}

// Gather into the (instance, Morton) order; sorted instance ids are the sort keys of the last pass.
__global__ __launch_bounds__(256) void k_gather_seg(uint32_t n, const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ sinst,
                                                    const uint32_t* __restrict__ mcode,
                                                    const float4* __restrict__ posh, const float* __restrict__ mass,
                                                    const float4* __restrict__ velu, const uint32_t* __restrict__ mat,
                                                    float4* __restrict__ spos, float* __restrict__ sh,
                                                    float4* __restrict__ svel, uint32_t* __restrict__ smat,
                                                    uint64_t* __restrict__ lkey) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t o = order[t];
    float4 p = posh[o];
    spos[t] = make_float4(p.x, p.y, p.z, mass[o]);
    if (sh) sh[t] = p.w;
    if (svel) svel[t] = velu[o];
    if (smat) smat[t] = mat[o];
    lkey[t] = ((uint64_t)sinst[t] << 30) | (uint64_t)mcode[o];
}

// Instance roots: the node whose leaf range is exactly the instance's (leaf size 1, so leaves are sorted
// particles and instance k's leaves are [ioff[k], ioff[k+1])).
__global__ __launch_bounds__(256) void k_inst_roots(uint32_t L, uint32_t ninst, const uint2* __restrict__ nrange,
                                                    const uint32_t* __restrict__ sinst,
                                                    const uint32_t* __restrict__ ioff, uint32_t* __restrict__ iroot) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < ninst) {  // one-particle instances are a leaf
        if (ioff[v + 1] - ioff[v] == 1) iroot[v] = (L - 1) + ioff[v];
    }
    if (v + 1 >= L) return;
    const uint2 r = nrange[v];
    const uint32_t k = sinst[r.x];
    if (r.x == ioff[k] && r.y + 1 == ioff[k + 1]) iroot[k] = v;
}

__global__ void k_reset_seg(uint32_t ninst, uint32_t* ibox, uint32_t* isig, int mode) {
    for (uint32_t k = threadIdx.x; k < ninst; k += blockDim.x) {
        isig[k] = 0x7f800000u;
        if (mode == 0) {
            for (int c = 0; c < 3; ++c) {
                ibox[6 * k + c] = 0xffffffffu;
                ibox[6 * k + 3 + c] = 0u;
            }
        }
    }
}

__global__ void k_set_dt_cfl_seg(uint32_t ninst, const uint32_t* __restrict__ isig, float cfl, float* idt) {
    for (uint32_t k = threadIdx.x; k < ninst; k += blockDim.x) {
        float m = __uint_as_float(isig[k]);
        idt[k] = (m < INFINITY) ? cfl * m : 0.0f;
    }
}
__global__ void k_set_dt_value_seg(uint32_t ninst, float dt, float* idt) {
    for (uint32_t k = threadIdx.x; k < ninst; k += blockDim.x) idt[k] = dt;
}
__global__ void k_advance_time_seg(uint32_t ninst, const float* __restrict__ idt, double* itime) {
    for (uint32_t k = threadIdx.x; k < ninst; k += blockDim.x) itime[k] += (double)idt[k];
}

void launch_reset_seg(const Dev& d, hipStream_t s, int mode) {
    hipLaunchKernelGGL(k_reset_seg, dim3(1), dim3(256), 0, s, d.ninst, d.ibox, d.isig, mode);
}
void launch_sorted_seg(const Dev& d, const float4* pos, const float* mass, hipStream_t s) {
    const uint32_t n = d.n;
    hipLaunchKernelGGL(k_bbox_seg, dim3(cdiv_e(n, 256)), dim3(256), 0, s, pos, n, d.inst, d.ibox);
    hipLaunchKernelGGL(k_morton_seg, dim3(cdiv_e(n, 256)), dim3(256), 0, s, pos, n, d.inst, d.ibox, d.keys, d.vals,
                       d.mcode);
    launch_radix_sort(d, s);  // (code, index) order, result in keys / vals
    hipLaunchKernelGGL(k_inst_key, dim3(cdiv_e(n, 256)), dim3(256), 0, s, n, d.vals, d.inst, d.keys_alt, d.vals_alt);
    launch_radix_pass(d, d.keys_alt, d.vals_alt, d.keys, d.vals, 0, s);  // stable by instance
    hipLaunchKernelGGL(k_gather_seg, dim3(cdiv_e(n, 256)), dim3(256), 0, s, n, d.vals, d.keys, d.mcode, pos, mass,
                       d.velu, d.mat, d.spos, d.sh, d.svel, d.smat, d.lkey);
}
void launch_inst_roots(const Dev& d, hipStream_t s) {
    const uint32_t threads = std::max(d.L, d.ninst);
    hipLaunchKernelGGL(k_inst_roots, dim3(cdiv_e(threads, 256)), dim3(256), 0, s, d.L, d.ninst, d.nrange, d.keys,
                       d.ioff, d.iroot);
}
void launch_set_dt_cfl_seg(const Dev& d, float cfl, hipStream_t s) {
    hipLaunchKernelGGL(k_set_dt_cfl_seg, dim3(1), dim3(256), 0, s, d.ninst, d.isig, cfl, d.idt);
}
void launch_set_dt_value_seg(const Dev& d, float dt, hipStream_t s) {
    hipLaunchKernelGGL(k_set_dt_value_seg, dim3(1), dim3(256), 0, s, d.ninst, dt, d.idt);
}
void launch_advance_time_seg(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_advance_time_seg, dim3(1), dim3(256), 0, s, d.ninst, d.idt, d.itime);
}

}  // namespace isph
