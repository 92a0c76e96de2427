// Bounding box, 30-bit Morton codes, LSD radix sort of (code, index) and the gather into Morton order.
//
// Reference: bh_gravity.wgsl:49-108 (bbox + Morton) and the CPU sort leg gpu_gravity.rs:351-388, which
// this replaces with a device radix sort. The sort is stable and its input is in index order, so it
// produces exactly the (code, index) order of `pairs.sort_unstable()` on (u32, u32) tuples.
#include "state.h"

namespace isph {

constexpr int kSortThreads = 256;
constexpr int kSortItems = kSortTile / kSortThreads;  // keys per thread per pass

// mode 0: everything before a tree build; mode 1: only what a force evaluation on the existing tree uses
// (Courant minimum, force-walk queue heads).
__global__ __launch_bounds__(1024) void k_reset_scalars(uint32_t* scal, int mode) {
    for (int i = threadIdx.x; i < kScalars; i += blockDim.x) {
        if (mode == 1) {
            if (i == 6) scal[i] = 0x7f800000u;
            if (i >= kQueueForce && i < kQueueForce + kQueueArea) scal[i] = 0u;
            continue;
        }
        if (i < 3) scal[i] = 0xffffffffu;        // min slots start at +inf key (bh_gravity.wgsl:64-65)
        else if (i < 6) scal[i] = 0u;            // max slots at -inf key
        else if (i == 6) scal[i] = 0x7f800000u;  // Courant signal min: +inf bits
        else scal[i] = 0u;                       // crossing nodes, group counts, work-queue areas
    }
}

// Block-reduced float-radix min/max (bh_gravity.wgsl:67-78 does one atomic per particle per axis).
__global__ __launch_bounds__(256) void k_bbox(const float4* __restrict__ pos, uint32_t n, uint32_t* scal) {
    __shared__ uint32_t red[6][4];
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
#pragma unroll 4
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = pos[i];
        uint32_t k0 = bbox_key(p.x), k1 = bbox_key(p.y), k2 = bbox_key(p.z);
        lo[0] = min(lo[0], k0); lo[1] = min(lo[1], k1); lo[2] = min(lo[2], k2);
        hi[0] = max(hi[0], k0); hi[1] = max(hi[1], k1); hi[2] = max(hi[2], k2);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = min(lo[k], (uint32_t)__shfl_xor((int)lo[k], o));
            hi[k] = max(hi[k], (uint32_t)__shfl_xor((int)hi[k], o));
        }
    }
    int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        for (int k = 0; k < 3; ++k) {
            red[k][w] = lo[k];
            red[3 + k][w] = hi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        uint32_t v = red[k][0];
        for (int j = 1; j < (int)(blockDim.x / kWave); ++j) v = (k < 3) ? min(v, red[k][j]) : max(v, red[k][j]);
        if (k < 3) atomicMin(&scal[k], v);
        else atomicMax(&scal[k], v);
    }
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {  // bh_gravity.wgsl:82-89
    v &= 0x000003ffu;
    v = (v * 0x00010001u) & 0xff0000ffu;
    v = (v * 0x00000101u) & 0x0f00f00fu;
    v = (v * 0x00000011u) & 0xc30c30c3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// bh_gravity.wgsl:91-108: clamp((p - lo) / ext) with correctly rounded f32 division, floor(u * 1024).
__global__ __launch_bounds__(256) void k_morton(const float4* __restrict__ pos, uint32_t n, const uint32_t* scal,
                                                uint32_t* __restrict__ codes, uint32_t* __restrict__ order) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lo[3], ext[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = bbox_decode(scal[k]);
        float hi = bbox_decode(scal[3 + k]);
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
    codes[i] = expand_bits(q[0]) * 4u + expand_bits(q[1]) * 2u + expand_bits(q[2]);
    order[i] = i;
}

// ---- LSD radix sort, 8-bit digits ------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                             int shift, uint32_t* __restrict__ hist, uint32_t nb) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
#pragma unroll 4
    for (int it = 0; it < kSortItems; ++it) {
        uint32_t e = base + it * kSortThreads + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of `total` counts in place by one block of 1024 threads, 16384 entries per tile: each
// thread takes 16 consecutive entries (four 16-byte loads when the run is whole), the run sums are scanned
// within each wave by shuffles and across the 16 waves through LDS, then each thread writes its run.
__global__ __launch_bounds__(1024) void k_radix_scan(uint32_t* data, uint32_t total) {
    constexpr uint32_t kPer = 16, kTile = 1024 * kPer, kWaves = 1024 / kWave;
    __shared__ uint32_t wsum[kWaves + 1];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < total; base += kTile) {
        const uint32_t i0 = base + threadIdx.x * kPer;
        uint32_t v[kPer];
        if (i0 + kPer <= total) {
            const uint4* q = reinterpret_cast<const uint4*>(data + i0);
#pragma unroll
            for (uint32_t k = 0; k < kPer / 4; ++k) {
                const uint4 x = q[k];
                v[4 * k] = x.x;
                v[4 * k + 1] = x.y;
                v[4 * k + 2] = x.z;
                v[4 * k + 3] = x.w;
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k) v[k] = i0 + k < total ? data[i0 + k] : 0u;
        }
        uint32_t s = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) s += v[k];
        uint32_t incl = s;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        if (lane == kWave - 1) wsum[w] = incl;
        __syncthreads();
        if (w == 0) {
            const uint32_t x = lane < (int)kWaves ? wsum[lane] : 0u;
            uint32_t xi = x;
#pragma unroll
            for (int o = 1; o < (int)kWaves; o <<= 1) {
                const uint32_t t = __shfl_up(xi, o);
                if (lane >= o) xi += t;
            }
            if (lane < (int)kWaves) wsum[lane] = xi - x;
            if (lane == (int)kWaves - 1) wsum[kWaves] = xi;
        }
        __syncthreads();
        uint32_t run = carry + wsum[w] + incl - s;
        if (i0 + kPer <= total) {
            uint4* q = reinterpret_cast<uint4*>(data + i0);
#pragma unroll
            for (uint32_t k = 0; k < kPer / 4; ++k) {
                uint4 x;
                x.x = run;
                run += v[4 * k];
                x.y = run;
                run += v[4 * k + 1];
                x.z = run;
                run += v[4 * k + 2];
                x.w = run;
                run += v[4 * k + 3];
                q[k] = x;
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k) {
                if (i0 + k < total) data[i0 + k] = run;
                run += v[k];
            }
        }
        carry += wsum[kWaves];
        __syncthreads();
    }
}

// Digit-major histogram scan for the sort passes: one wave per digit scans that digit's per-block counts
// into exclusive offsets within the digit (4 consecutive counts per lane, a wave shuffle scan) and stores
// the digit total after the histogram (hist[256 * nb + digit]); each scatter block adds the digit bases.
__global__ __launch_bounds__(256) void k_radix_colscan(uint32_t* hist, uint32_t nb) {
    const uint32_t d = blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    uint32_t* col = hist + (size_t)d * nb;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 4 * kWave) {
        uint32_t v[4], s = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t idx = base + 4 * lane + k;
            v[k] = idx < nb ? col[idx] : 0u;
            s += v[k];
        }
        uint32_t incl = s;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        uint32_t run = carry + incl - s;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t idx = base + 4 * lane + k;
            if (idx < nb) col[idx] = run;
            run += v[k];
        }
        carry += __shfl(incl, kWave - 1);
    }
    if (lane == 0) hist[256u * nb + d] = carry;
}

// Stable scatter: within a tile, elements are ranked in index order per digit with wave64 match masks.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout,
                                                                uint32_t* __restrict__ vout, uint32_t n, int shift,
                                                                const uint32_t* __restrict__ hist, uint32_t nb) {
    constexpr int kWaves = kSortThreads / kWave;
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kWaves][256];
    int tid = threadIdx.x, w = tid / kWave;
    // digit bases: exclusive scan of the digit totals
    const uint32_t tot = hist[256u * nb + tid];
    run[tid] = tot;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t t = tid >= o ? run[tid - o] : 0u;
        __syncthreads();
        run[tid] += t;
        __syncthreads();
    }
    const uint32_t dbase = run[tid] - tot;
    __syncthreads();
    run[tid] = dbase + hist[tid * nb + blockIdx.x];
#pragma unroll
    for (int k = 0; k < kWaves; ++k) wcnt[k][tid] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
        uint32_t e = base + it * kSortThreads + tid;
        bool valid = e < n;
        uint32_t key = valid ? kin[e] : 0xffffffffu;
        uint32_t val = valid ? vin[e] : 0u;
        uint32_t dig = (key >> shift) & 0xffu;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bal = __ballot((dig >> b) & 1u);
            peers &= ((dig >> b) & 1u) ? bal : ~bal;
        }
        int rank = __popcll(peers & lanemask_lt());
        if (rank == 0) wcnt[w][dig] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t off = run[dig] + rank;
        for (int k = 0; k < w; ++k) off += wcnt[k][dig];
        if (valid) {
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            add += wcnt[k][tid];
            wcnt[k][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
}

// Gather the state into Morton order; leaf cluster code = first code of each bucket (gpu_gravity.rs:386).
__global__ __launch_bounds__(256) void k_gather(uint32_t n, uint32_t K, const uint32_t* __restrict__ order,
                                                const uint32_t* __restrict__ codes, const float4* __restrict__ posh,
                                                const float* __restrict__ mass, const float4* __restrict__ velu,
                                                const uint32_t* __restrict__ mat, float4* __restrict__ spos,
                                                float* __restrict__ sh, float4* __restrict__ svel,
                                                uint32_t* __restrict__ smat, uint32_t* __restrict__ lcode) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t o = order[t];
    float4 p = posh[o];
    spos[t] = make_float4(p.x, p.y, p.z, mass[o]);
    if (sh) sh[t] = p.w;
    if (svel) svel[t] = velu[o];
    if (smat) smat[t] = mat[o];
    if (t % K == 0) lcode[t / K] = codes[t];
}

__global__ __launch_bounds__(256) void k_gather_vel(uint32_t n, const uint32_t* __restrict__ order,
                                                    const float4* __restrict__ velu, float4* __restrict__ svel) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    svel[t] = velu[order[t]];
}

static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

void launch_reset_scalars(const Dev& d, hipStream_t s, int mode) {
    hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(1024), 0, s, d.scal, mode);
}
void launch_bbox(const Dev& d, const float4* pos, hipStream_t s) {
    // few blocks: each ends in six same-address atomics, which serialise in L2
    uint32_t blocks = std::min<uint32_t>(cdiv(d.n, 256), 256);
    hipLaunchKernelGGL(k_bbox, dim3(std::max<uint32_t>(blocks, 1)), dim3(256), 0, s, pos, d.n, d.scal);
}
void launch_morton(const Dev& d, const float4* pos, hipStream_t s) {
    hipLaunchKernelGGL(k_morton, dim3(cdiv(d.n, 256)), dim3(256), 0, s, pos, d.n, d.scal, d.keys, d.vals);
}
void launch_radix_sort(const Dev& d, hipStream_t s) {
    uint32_t nb = cdiv(d.n, kSortTile);
    uint32_t* kin = d.keys;
    uint32_t* vin = d.vals;
    uint32_t* kout = d.keys_alt;
    uint32_t* vout = d.vals_alt;
    for (int shift = 0; shift < 32; shift += 8) {  // 4 passes: even count, result back in keys/vals
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, s, kin, d.n, shift, d.hist, nb);
        hipLaunchKernelGGL(k_radix_colscan, dim3(256 / (256 / kWave)), dim3(256), 0, s, d.hist, nb);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, s, kin, vin, kout, vout, d.n, shift,
                           d.hist, nb);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
}
void launch_radix_pass(const Dev& d, const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                       int shift, hipStream_t s) {
    uint32_t nb = cdiv(d.n, kSortTile);
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, s, kin, d.n, shift, d.hist, nb);
    hipLaunchKernelGGL(k_radix_colscan, dim3(256 / (256 / kWave)), dim3(256), 0, s, d.hist, nb);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, s, kin, vin, kout, vout, d.n, shift, d.hist,
                       nb);
}
void launch_radix_pass_n(uint32_t n, uint32_t* hist, const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                         uint32_t* vout, int shift, hipStream_t s) {
    if (n == 0) return;
    uint32_t nb = cdiv(n, kSortTile);
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, s, kin, n, shift, hist, nb);
    hipLaunchKernelGGL(k_radix_colscan, dim3(256 / (256 / kWave)), dim3(256), 0, s, hist, nb);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, s, kin, vin, kout, vout, n, shift, hist, nb);
}
void launch_exclusive_scan(uint32_t* data, uint32_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, data, count);
}
void launch_gather(const Dev& d, const float4* pos, const float* mass, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3(cdiv(d.n, 256)), dim3(256), 0, s, d.n, d.K, d.vals, d.keys, pos, mass, d.velu,
                       d.mat, d.spos, d.sh, d.svel, d.smat, d.lcode);
}
void launch_gather_vel(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_vel, dim3(cdiv(d.n, 256)), dim3(256), 0, s, d.n, d.vals, d.velu, d.svel);
}

}  // namespace isph
