// On-device measurement: the disk view (gpu_sph.rs:689-785, orbit.rs:91-117) and the friends-of-friends
// clump finder (accretion.rs:158-301), so an ensemble is measured without reading its state back.
//
// Disk view, per instance:
//   live sums        chunk partials -> COM, bulk velocity, total mass, mean h (gpu_sph.rs:697-719, :767)
//   remnant          radius sort of the live particles about the COM (f64 radius bits as the key) and the
//                    enclosed-mass scan: the first radius where the mass reaches 85 % (:722-736)
//   classification   is_bound first, then perigee > r_remnant -> disk (:739-765); remnant and escaped
//                    mass split by source body as impact-run's measure does (main.rs:752-760)
//   members          the disk particles gathered in upload order, in f64, with rho.max(1) (:760)
// Clumps, per instance (accretion.rs:176-301):
//   cells            hash of the linking-length cell (one part in 1e9 wider, so a linked pair is always
//                    in neighbouring cells), members sorted by it
//   union-find       lock-free: the larger root is hooked under the smaller with a compare-and-swap, so
//                    each component ends rooted at its smallest member (the clump id)
//   statistics       one wave per clump: mass, COM, bulk velocity, volume density, internal KE, spin,
//                    inertia tensor -> coherent-rotation energy and heat
//   self energy      exact softened pair sum over 256 x 256 tiles of each clump, summed per clump in tile
//                    order (the reference is exact below 1024 members and takes a theta = 0.5
//                    octree above)
//   gates            bound = KE + PE < 0; outside the fluid Roche limit of the remnant
// The orbit and distance decisions run with floating-point contraction off, in the reference's operation
// order, so they round as the reference does.
#include "measure.h"

namespace isph {

namespace {

constexpr int kMT = 256;                                           // threads per chunk block
constexpr double kFourThirdsPi = 4.0 / 3.0 * 3.14159265358979323846;  // accretion.rs:31
constexpr uint64_t kInfBits = 0x7ff0000000000000ull;
constexpr int kUnionIters = 1 << 16;
constexpr double kCellPad = 1.0 + 1.0e-9;
constexpr double kCellClamp = 1048576.0;

inline uint32_t cdiv_m(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

struct D3 {
    double x, y, z;
};

__device__ __forceinline__ double dot3(D3 a, D3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ D3 cross3(D3 a, D3 b) {  // glam DVec3::cross
#pragma clang fp contract(off)
    return D3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

__device__ __forceinline__ bool live_of(float4 p, float4 v, float m) {  // gpu_sph.rs:700-707
    return m > 0.0f && isfinite(m) && isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(v.x) &&
           isfinite(v.y) && isfinite(v.z);
}

// orbit.rs:91-94
__device__ __forceinline__ bool is_bound(D3 p, D3 v, double mu) {
#pragma clang fp contract(off)
    const double r = sqrt(dot3(p, p));
    return r > 0.0 && 0.5 * dot3(v, v) - mu / r < 0.0;
}
// orbit.rs:96-117 for a bound orbit (energy < 0, r > 0): a (1 - e).
__device__ __forceinline__ double perigee_bound(D3 p, D3 v, double mu) {
#pragma clang fp contract(off)
    const double r = sqrt(dot3(p, p));
    const double energy = 0.5 * dot3(v, v) - mu / r;
    const D3 hc = cross3(p, v);
    const double h = sqrt(dot3(hc, hc));
    const double e = sqrt(fmax(1.0 + 2.0 * energy * h * h / (mu * mu), 0.0));
    const double a = -mu / (2.0 * energy);
    return a * (1.0 - e);
}

__device__ __forceinline__ double block_sum(double* red, double v) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kMT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Radius about the instance COM of a live particle of a valid instance, else -1.
__device__ __forceinline__ double radius_of(uint32_t i, const float4* posh, const float4* velu, const float* mass,
                                            const double* vw) {
#pragma clang fp contract(off)
    if (vw[kVValid] == 0.0) return -1.0;
    const float4 p = posh[i], v = velu[i];
    const float m = mass[i];
    if (!live_of(p, v, m)) return -1.0;
    const D3 d{(double)p.x - vw[kVCom], (double)p.y - vw[kVCom + 1], (double)p.z - vw[kVCom + 2]};
    return sqrt(dot3(d, d));
}

__device__ __forceinline__ uint32_t chunk_of(uint32_t i, uint32_t k, const uint32_t* ioff, const uint32_t* ichunk) {
    const uint32_t start = ioff ? ioff[k] : 0u;
    return ichunk[k] + (i - start) / kMeasChunk;
}

// ---- live sums ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kMT) void k_meas_live(const uint2* __restrict__ chunks, const float4* __restrict__ posh,
                                                   const float4* __restrict__ velu, const float* __restrict__ mass,
                                                   double* __restrict__ part) {
    __shared__ double red[kMT];
    const uint2 r = chunks[blockIdx.x];
    double f[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k] = 0.0;
    for (int q = 0; q < 4; ++q) {
        const uint32_t i = r.x + threadIdx.x * 4 + q;
        if (i >= r.y) break;
        const float4 p = posh[i], v = velu[i];
        const float m = mass[i];
        if (!live_of(p, v, m)) continue;
        const double md = m;
        f[0] += 1.0;
        f[1] += md;
        f[2] += (double)p.x * md;
        f[3] += (double)p.y * md;
        f[4] += (double)p.z * md;
        f[5] += (double)v.x * md;
        f[6] += (double)v.y * md;
        f[7] += (double)v.z * md;
        f[8] += (double)p.w;
    }
    for (int k = 0; k < 9; ++k) {
        const double s = block_sum(red, f[k]);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * kPartFields + k] = s;
    }
}

// One block per instance: chunk partials summed in a fixed order (strided per thread, then a tree).
__global__ __launch_bounds__(kMT) void k_meas_view1(const uint32_t* __restrict__ ichunk,
                                                    const double* __restrict__ part, double* __restrict__ view,
                                                    uint32_t* __restrict__ cross) {
    __shared__ double red[kMT];
    const uint32_t k = blockIdx.x;
    double f[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) f[j] = 0.0;
    for (uint32_t c = ichunk[k] + threadIdx.x; c < ichunk[k + 1]; c += kMT)
#pragma unroll
        for (int j = 0; j < 9; ++j) f[j] += part[(size_t)c * kPartFields + j];
#pragma unroll
    for (int j = 0; j < 9; ++j) f[j] = block_sum(red, f[j]);
    if (threadIdx.x != 0) return;
    double* vw = view + (size_t)k * kViewFields;
    for (int j = 0; j < kViewFields; ++j) vw[j] = 0.0;
    const double n_live = f[0], mt = f[1];
    vw[kVNLive] = n_live;
    vw[kVMTot] = mt;
    if (n_live >= 2.0 && mt > 0.0) {  // gpu_sph.rs:708-715
        vw[kVValid] = 1.0;
        for (int a = 0; a < 3; ++a) {
            vw[kVCom + a] = f[2 + a] / mt;
            vw[kVVCom + a] = f[5 + a] / mt;
        }
        vw[kVMeanH] = f[8] / n_live;
    }
    cross[k] = 0xffffffffu;
}

// ---- remnant: radius sort and enclosed-mass scan -------------------------------------------------------
__global__ __launch_bounds__(256) void k_meas_rkey(uint32_t n, const uint32_t* __restrict__ inst,
                                                   const float4* __restrict__ posh, const float4* __restrict__ velu,
                                                   const float* __restrict__ mass, const double* __restrict__ view,
                                                   uint32_t* __restrict__ klo, uint32_t* __restrict__ khi,
                                                   uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = inst ? inst[i] : 0u;
    const double r = radius_of(i, posh, velu, mass, view + (size_t)k * kViewFields);
    const uint64_t bits = r >= 0.0 ? (uint64_t)__double_as_longlong(r) : kInfBits;  // r >= 0: bits are monotone
    klo[i] = (uint32_t)bits;
    khi[i] = (uint32_t)(bits >> 32);
    vals[i] = i;
}

__global__ __launch_bounds__(256) void k_gather_u32(uint32_t n, const uint32_t* __restrict__ src,
                                                    const uint32_t* __restrict__ idx, uint32_t* __restrict__ dst) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[idx[p]];
}

// Enclosed mass in radius order: 4 consecutive positions per thread, thread totals scanned in order by one
// thread, so the cumulative mass is non-decreasing along the sorted order (each chunk total is exactly the
// cumulative value of its last position).
__global__ __launch_bounds__(kMT) void k_meas_scan(const uint2* __restrict__ chunks, const uint32_t* __restrict__ cinst,
                                                   const uint32_t* __restrict__ vs, const float4* __restrict__ posh,
                                                   const float4* __restrict__ velu, const float* __restrict__ mass,
                                                   const double* __restrict__ view, double* __restrict__ scum,
                                                   double* __restrict__ ctot) {
    __shared__ double tot[kMT];
    const uint2 r = chunks[blockIdx.x];
    const double* vw = view + (size_t)cinst[blockIdx.x] * kViewFields;
    double loc[4], t = 0.0;
    for (int q = 0; q < 4; ++q) {
        const uint32_t p = r.x + threadIdx.x * 4 + q;
        double m = 0.0;
        if (p < r.y) {
            const uint32_t i = vs[p];
            if (radius_of(i, posh, velu, mass, vw) >= 0.0) m = (double)mass[i];
        }
        t += m;
        loc[q] = t;
    }
    tot[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int j = 0; j < kMT; ++j) {
            const double x = tot[j];
            tot[j] = run;
            run += x;
        }
        ctot[blockIdx.x] = run;
    }
    __syncthreads();
    const double ex = tot[threadIdx.x];
    for (int q = 0; q < 4; ++q) {
        const uint32_t p = r.x + threadIdx.x * 4 + q;
        if (p < r.y) scum[p] = ex + loc[q];
    }
}

// Exclusive scan of one instance's chunk totals (one block per instance), left to right in one thread so
// that each chunk's offset is exactly the previous offset plus the previous total: the cumulative mass
// stays non-decreasing across chunk boundaries. The totals are staged through LDS so the serial chain
// runs on LDS reads rather than on global-memory latency.
constexpr uint32_t kCoffTile = 4096;
__global__ __launch_bounds__(kMT) void k_meas_coff(const uint32_t* __restrict__ ichunk, double* __restrict__ coff) {
    __shared__ double buf[kCoffTile];
    const uint32_t c0 = ichunk[blockIdx.x], c1 = ichunk[blockIdx.x + 1];
    double run = 0.0;
    for (uint32_t base = c0; base < c1; base += kCoffTile) {
        const uint32_t len = min(kCoffTile, c1 - base);
        for (uint32_t j = threadIdx.x; j < len; j += kMT) buf[j] = coff[base + j];
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t j = 0;
            for (; j + 8 <= len; j += 8) {
                double x[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) x[q] = buf[j + q];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    buf[j + q] = run;
                    run += x[q];
                }
            }
            for (; j < len; ++j) {
                const double x = buf[j];
                buf[j] = run;
                run += x;
            }
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < len; j += kMT) coff[base + j] = buf[j];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_meas_cross(uint32_t n, const uint32_t* __restrict__ inst,
                                                    const uint32_t* __restrict__ ioff,
                                                    const uint32_t* __restrict__ ichunk,
                                                    const uint32_t* __restrict__ vs, const float4* __restrict__ posh,
                                                    const float4* __restrict__ velu, const float* __restrict__ mass,
                                                    const double* __restrict__ view, const double* __restrict__ scum,
                                                    const double* __restrict__ coff, uint32_t* __restrict__ cross) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = inst ? inst[p] : 0u;  // instance segments coincide in upload and sorted order
    const double* vw = view + (size_t)k * kViewFields;
    if (radius_of(vs[p], posh, velu, mass, vw) < 0.0) return;
    const uint32_t start = ioff ? ioff[k] : 0u;
    const double cum = coff[chunk_of(p, k, ioff, ichunk)] + scum[p];
    const double prev = p == start ? 0.0 : coff[chunk_of(p - 1, k, ioff, ichunk)] + scum[p - 1];
    const double T = 0.85 * vw[kVMTot];  // gpu_sph.rs:728-735
    if (cum >= T && prev < T) cross[k] = p;
}

__global__ void k_meas_view2(uint32_t ninst, const uint32_t* __restrict__ ioff, const uint32_t* __restrict__ ichunk,
                             const uint32_t* __restrict__ vs, const float4* __restrict__ posh,
                             const float4* __restrict__ velu, const float* __restrict__ mass,
                             const double* __restrict__ scum, const double* __restrict__ coff,
                             const uint32_t* __restrict__ cross, double* __restrict__ view) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    double* vw = view + (size_t)k * kViewFields;
    if (vw[kVValid] == 0.0) return;
    const uint32_t p = cross[k];
    double r, mr;
    if (p != 0xffffffffu) {
        r = radius_of(vs[p], posh, velu, mass, vw);
        mr = coff[chunk_of(p, k, ioff, ichunk)] + scum[p];
    } else {  // no crossing: the largest radius and the whole mass (gpu_sph.rs:726-727)
        const uint32_t last = (ioff ? ioff[k] : 0u) + (uint32_t)vw[kVNLive] - 1u;
        r = radius_of(vs[last], posh, velu, mass, vw);
        mr = vw[kVMTot];
    }
    vw[kVRRem] = r;
    vw[kVMRem] = mr;
    vw[kVMu] = kG64 * mr;  // gpu_sph.rs:737
}

// ---- classification -------------------------------------------------------------------------------------
__global__ __launch_bounds__(kMT) void k_meas_classify(const uint2* __restrict__ chunks,
                                                       const uint32_t* __restrict__ cinst,
                                                       const float4* __restrict__ posh,
                                                       const float4* __restrict__ velu,
                                                       const float* __restrict__ mass,
                                                       const uint32_t* __restrict__ prov,
                                                       const double* __restrict__ view, double* __restrict__ part,
                                                       uint32_t* __restrict__ slot) {
#pragma clang fp contract(off)
    __shared__ double red[kMT];
    __shared__ uint32_t cnt[kMT];
    const uint2 r = chunks[blockIdx.x];
    const double* vw = view + (size_t)cinst[blockIdx.x] * kViewFields;
    const bool valid = vw[kVValid] != 0.0;
    const D3 com{vw[kVCom], vw[kVCom + 1], vw[kVCom + 2]};
    const D3 vcom{vw[kVVCom], vw[kVVCom + 1], vw[kVVCom + 2]};
    const double mu = vw[kVMu], r_rem = vw[kVRRem];
    // disk prov 0, disk other, escaped, remnant prov 0, remnant other, escaped prov 0
    double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    uint32_t flags = 0, c = 0;
    for (int q = 0; q < 4; ++q) {
        const uint32_t i = r.x + threadIdx.x * 4 + q;
        if (i >= r.y || !valid) continue;
        const float4 p = posh[i], v = velu[i];
        const float mf = mass[i];
        if (!live_of(p, v, mf)) continue;
        const double m = mf;
        const bool p0 = prov[i] == 0u;
        const D3 rp{(double)p.x - com.x, (double)p.y - com.y, (double)p.z - com.z};
        const D3 rv{(double)v.x - vcom.x, (double)v.y - vcom.y, (double)v.z - vcom.z};
        if (!is_bound(rp, rv, mu)) {  // bound first, then geometry (gpu_sph.rs:745-750)
            f[2] += m;
            if (p0) f[5] += m;
            continue;
        }
        if (perigee_bound(rp, rv, mu) > r_rem) {
            f[p0 ? 0 : 1] += m;
            flags |= 1u << q;
            ++c;
        } else {
            f[p0 ? 3 : 4] += m;
        }
    }
    for (int k = 0; k < 6; ++k) {
        const double s = block_sum(red, f[k]);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * kPartFields + 9 + k] = s;
    }
    // slots of the disk particles within the chunk, in upload order
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int o = 1; o < kMT; o <<= 1) {
        const uint32_t t = (int)threadIdx.x >= o ? cnt[threadIdx.x - o] : 0u;
        __syncthreads();
        cnt[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t s = cnt[threadIdx.x] - c;
    if (threadIdx.x == kMT - 1) part[(size_t)blockIdx.x * kPartFields + 15] = (double)cnt[kMT - 1];
    for (int q = 0; q < 4; ++q) {
        const uint32_t i = r.x + threadIdx.x * 4 + q;
        if (i >= r.y) break;
        slot[i] = (flags >> q) & 1u ? s++ : 0xffffffffu;
    }
}

// Classification sums per instance (one block per instance, fixed order) and its clump-finder parameters.
__global__ __launch_bounds__(kMT) void k_meas_view3(const uint32_t* __restrict__ ichunk,
                                                    const double* __restrict__ part, double* __restrict__ view,
                                                    double* __restrict__ iparam, double clump_soft,
                                                    double link_factor) {
    __shared__ double red[kMT];
    const uint32_t k = blockIdx.x;
    double f[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) f[j] = 0.0;
    for (uint32_t c = ichunk[k] + threadIdx.x; c < ichunk[k + 1]; c += kMT)
#pragma unroll
        for (int j = 0; j < 7; ++j) f[j] += part[(size_t)c * kPartFields + 9 + j];
#pragma unroll
    for (int j = 0; j < 7; ++j) f[j] = block_sum(red, f[j]);
    if (threadIdx.x != 0) return;
    double* vw = view + (size_t)k * kViewFields;
    vw[kVDisk0] = f[0];
    vw[kVDiskO] = f[1];
    vw[kVEsc] = f[2];
    vw[kVRem0] = f[3];
    vw[kVRemO] = f[4];
    vw[kVEsc0] = f[5];
    vw[kVNDisk] = f[6];
    double* pr = iparam + (size_t)k * kParamFields;
    pr[kPLink] = link_factor * vw[kVMeanH];  // gpu_sph.rs:803: 2 x mean h of the whole field
    pr[kPSoft2] = clump_soft * clump_soft;
    pr[kPG] = kG64;
    for (int a = 0; a < 3; ++a) pr[kPCPos + a] = vw[kVCom + a];
    pr[kPCMass] = vw[kVMRem];
    pr[kPCRad] = vw[kVRRem];
}

// Member base of every chunk (exclusive scan of the per-chunk disk counts, integers, so the order of the
// additions does not matter), the first member of each instance and the member count. One block of
// 1024 threads, 1024 chunks per tile, wave scans plus a carry.
constexpr int kMbT = 1024;
__global__ __launch_bounds__(kMbT) void k_meas_mbase(uint32_t ninst, uint32_t nchunks,
                                                     const uint32_t* __restrict__ ichunk,
                                                     const double* __restrict__ part, uint32_t* __restrict__ mbase,
                                                     uint32_t* __restrict__ moff, uint32_t* __restrict__ word) {
    __shared__ uint32_t wsum[kMbT / 64];
    __shared__ uint32_t carry_s;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nchunks; base += kMbT) {
        const uint32_t c = base + threadIdx.x;
        const uint32_t x = c < nchunks ? (uint32_t)part[(size_t)c * kPartFields + 15] : 0u;
        uint32_t inc = x;  // inclusive wave scan
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o);
            if (lane >= o) inc += t;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t r = carry;
            for (int q = 0; q < kMbT / 64; ++q) {
                const uint32_t t = wsum[q];
                wsum[q] = r;
                r += t;
            }
            carry_s = r;
        }
        __syncthreads();
        if (c < nchunks) mbase[c] = wsum[w] + inc - x;
        carry = carry_s;
        __syncthreads();
    }
    for (uint32_t k = threadIdx.x; k < ninst; k += kMbT) moff[k] = ichunk[k] < nchunks ? mbase[ichunk[k]] : carry;
    if (threadIdx.x == 0) {
        moff[ninst] = carry;
        word[0] = carry;
    }
}

__global__ __launch_bounds__(256) void k_meas_gather(uint32_t n, const uint32_t* __restrict__ inst,
                                                     const uint32_t* __restrict__ ioff,
                                                     const uint32_t* __restrict__ ichunk,
                                                     const float4* __restrict__ posh, const float4* __restrict__ velu,
                                                     const float* __restrict__ mass, const float* __restrict__ rho,
                                                     const uint32_t* __restrict__ slot,
                                                     const uint32_t* __restrict__ mbase, double4* __restrict__ mp,
                                                     double4* __restrict__ mv, uint32_t* __restrict__ minst,
                                                     uint32_t* __restrict__ midx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = slot[i];
    if (s == 0xffffffffu) return;
    const uint32_t k = inst ? inst[i] : 0u;
    const uint32_t g = mbase[chunk_of(i, k, ioff, ichunk)] + s;
    const float4 p = posh[i], v = velu[i];
    mp[g] = make_double4(p.x, p.y, p.z, mass[i]);
    mv[g] = make_double4(v.x, v.y, v.z, (double)fmaxf(rho[i], 1.0f));  // gpu_sph.rs:760
    minst[g] = k;
    midx[g] = i - (ioff ? ioff[k] : 0u);
}

// ---- friends-of-friends ---------------------------------------------------------------------------------
__device__ __forceinline__ int cell_coord(double x, double o, double cs) {
    double q = floor((x - o) / cs);
    q = fmin(fmax(q, -kCellClamp), kCellClamp);
    return (int)q;
}
__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
__device__ __forceinline__ uint32_t cell_hash(uint32_t k, int cx, int cy, int cz) {
    uint32_t h = k * 0x9E3779B1u + 0x7F4A7C15u;
    h = rotl32(h ^ ((uint32_t)cx * 0x85EBCA77u), 13) * 0xC2B2AE3Du;
    h = rotl32(h ^ ((uint32_t)cy * 0x27D4EB2Fu), 13) * 0x165667B1u;
    h = rotl32(h ^ ((uint32_t)cz * 0x9E3779B1u), 13) * 0x85EBCA77u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ double cell_size(const double* pr) { return fmax(pr[kPLink], 1.0e-9) * kCellPad; }

__global__ __launch_bounds__(256) void k_fof_key(uint32_t n, const double4* __restrict__ mp,
                                                 const uint32_t* __restrict__ minst, const double* __restrict__ iparam,
                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                 uint32_t* __restrict__ parent) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t k = minst[s];
    const double* pr = iparam + (size_t)k * kParamFields;
    const double cs = cell_size(pr);
    const double4 p = mp[s];
    keys[s] = cell_hash(k, cell_coord(p.x, pr[kPCPos], cs), cell_coord(p.y, pr[kPCPos + 1], cs),
                        cell_coord(p.z, pr[kPCPos + 2], cs));
    vals[s] = s;
    parent[s] = s;
}

__device__ __forceinline__ uint32_t uf_load(uint32_t* a) {
    return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void uf_store(uint32_t* a, uint32_t v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Root of x with path halving. Every parent link points to a smaller index, so the walk ends.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    while (true) {
        const uint32_t p = uf_load(parent + x);
        if (p == x) return x;
        const uint32_t g = uf_load(parent + p);
        if (g == p) return p;
        uf_store(parent + x, g);
        x = g;
    }
}
__device__ __forceinline__ void uf_unite(uint32_t* parent, uint32_t a, uint32_t b, uint32_t* overflow) {
    for (int it = 0; it < kUnionIters; ++it) {
        uint32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
        if (ra == rb) return;
        if (ra < rb) {
            const uint32_t t = ra;
            ra = rb;
            rb = t;
        }
        if (atomicCAS(parent + ra, ra, rb) == ra) return;
        a = ra;
        b = rb;
    }
    atomicOr(overflow, 1u);
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* keys, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// accretion.rs:192-201: link every pair within the linking length (squared distances compared, f64).
__global__ __launch_bounds__(256) void k_fof_link(uint32_t n, const uint32_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ vals, const double4* __restrict__ mp,
                                                  const uint32_t* __restrict__ minst,
                                                  const double* __restrict__ iparam, uint32_t* parent,
                                                  uint32_t* overflow) {
#pragma clang fp contract(off)
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t s = vals[p];
    const uint32_t k = minst[s];
    const double* pr = iparam + (size_t)k * kParamFields;
    const double ll = pr[kPLink], ll2 = ll * ll;
    const double cs = cell_size(pr);
    const double4 a = mp[s];
    const int cx = cell_coord(a.x, pr[kPCPos], cs), cy = cell_coord(a.y, pr[kPCPos + 1], cs),
              cz = cell_coord(a.z, pr[kPCPos + 2], cs);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const uint32_t key = cell_hash(k, cx + dx, cy + dy, cz + dz);
                for (uint32_t q = lower_bound_u32(keys, n, key); q < n && keys[q] == key; ++q) {
                    const uint32_t t = vals[q];
                    if (t <= s || minst[t] != k) continue;
                    const double4 b = mp[t];
                    const D3 d{a.x - b.x, a.y - b.y, a.z - b.z};
                    if (dot3(d, d) <= ll2) uf_unite(parent, s, t, overflow);
                }
            }
}

__global__ __launch_bounds__(256) void k_fof_label(uint32_t n, uint32_t* parent, uint32_t* __restrict__ label,
                                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t r = uf_find(parent, s);
    label[s] = r;
    keys[s] = r;
    vals[s] = s;
}

__global__ __launch_bounds__(256) void k_clump_flag(uint32_t n, const uint32_t* __restrict__ ks,
                                                    uint32_t* __restrict__ flag) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) flag[p] = (p == 0 || ks[p] != ks[p - 1]) ? 1u : 0u;
    if (p == n) flag[p] = 0u;
}

__global__ __launch_bounds__(256) void k_clump_emit(uint32_t n, const uint32_t* __restrict__ ks,
                                                    const uint32_t* __restrict__ vs,
                                                    const uint32_t* __restrict__ scan, uint32_t* __restrict__ cstart,
                                                    const double4* __restrict__ mp, const double4* __restrict__ mv,
                                                    double4* __restrict__ cp, double4* __restrict__ cv,
                                                    uint32_t* __restrict__ word) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (p == 0 || ks[p] != ks[p - 1]) cstart[scan[p]] = p;
    const uint32_t s = vs[p];
    cp[p] = mp[s];
    cv[p] = mv[s];
    if (p == n - 1) {
        cstart[scan[n]] = n;
        word[8] = scan[n];
    }
}

// ---- per-clump statistics: one wave per clump ------------------------------------------------------------
__device__ __forceinline__ double wave_sum_lane0(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    return __shfl(v, 0);
}

__global__ __launch_bounds__(256) void k_clump_stats(uint32_t nc, const uint32_t* __restrict__ cstart,
                                                     const uint32_t* __restrict__ vs,
                                                     const uint32_t* __restrict__ minst,
                                                     const double4* __restrict__ cp, const double4* __restrict__ cv,
                                                     double* __restrict__ crec, uint32_t* __restrict__ tcnt) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const uint32_t w0 = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const uint32_t ws = gridDim.x * (blockDim.x / 64);
    for (uint32_t c = w0; c < nc; c += ws) {
        const uint32_t s0 = cstart[c], s1 = cstart[c + 1], cnt = s1 - s0;
        double m = 0, px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, vol = 0;
        for (uint32_t e = s0 + lane; e < s1; e += 64) {
            const double4 P = cp[e], V = cv[e];
            m += P.w;
            px += P.x * P.w;
            py += P.y * P.w;
            pz += P.z * P.w;
            vx += V.x * P.w;
            vy += V.y * P.w;
            vz += V.z * P.w;
            vol += P.w / V.w;
        }
        m = wave_sum_lane0(m);
        px = wave_sum_lane0(px);
        py = wave_sum_lane0(py);
        pz = wave_sum_lane0(pz);
        vx = wave_sum_lane0(vx);
        vy = wave_sum_lane0(vy);
        vz = wave_sum_lane0(vz);
        vol = wave_sum_lane0(vol);
        const D3 com{px / m, py / m, pz / m}, cvel{vx / m, vy / m, vz / m};
        const double rho = vol > 0.0 ? m / vol : cv[s0].w;  // accretion.rs:219-224
        const double radius = cbrt(m / (kFourThirdsPi * rho));
        double ke = 0, lx = 0, ly = 0, lz = 0, ixx = 0, iyy = 0, izz = 0, ixy = 0, ixz = 0, iyz = 0;
        for (uint32_t e = s0 + lane; e < s1; e += 64) {
            const double4 P = cp[e], V = cv[e];
            const D3 r{P.x - com.x, P.y - com.y, P.z - com.z};
            const D3 dv{V.x - cvel.x, V.y - cvel.y, V.z - cvel.z};
            ke += 0.5 * P.w * dot3(dv, dv);
            const D3 l = cross3(r, dv);
            lx += P.w * l.x;
            ly += P.w * l.y;
            lz += P.w * l.z;
            const double r2 = dot3(r, r), mr2 = P.w * r2;
            ixx += mr2 - r.x * (P.w * r.x);
            iyy += mr2 - r.y * (P.w * r.y);
            izz += mr2 - r.z * (P.w * r.z);
            ixy -= r.x * (P.w * r.y);
            ixz -= r.x * (P.w * r.z);
            iyz -= r.y * (P.w * r.z);
        }
        ke = wave_sum_lane0(ke);
        lx = wave_sum_lane0(lx);
        ly = wave_sum_lane0(ly);
        lz = wave_sum_lane0(lz);
        ixx = wave_sum_lane0(ixx);
        iyy = wave_sum_lane0(iyy);
        izz = wave_sum_lane0(izz);
        ixy = wave_sum_lane0(ixy);
        ixz = wave_sum_lane0(ixz);
        iyz = wave_sum_lane0(iyz);
        if (lane == 0) {
            // coherent-rotation energy 1/2 (I^-1 L).L (accretion.rs:243-260); columns of I as in glam's DMat3
            const D3 c0{ixx, ixy, ixz}, c1{ixy, iyy, iyz}, c2{ixz, iyz, izz};
            const D3 t0 = cross3(c1, c2), t1 = cross3(c2, c0), t2 = cross3(c0, c1);
            const double det = dot3(c2, t2);
            double e_rot = 0.0;
            if (cnt >= 3 && isfinite(det) && det > 0.0) {
                const double id = 1.0 / det;
                const D3 L{lx, ly, lz};
                const D3 w{t0.x * id * L.x + t0.y * id * L.y + t0.z * id * L.z,
                           t1.x * id * L.x + t1.y * id * L.y + t1.z * id * L.z,
                           t2.x * id * L.x + t2.y * id * L.y + t2.z * id * L.z};
                e_rot = 0.5 * dot3(w, L);
            }
            double* rec = crec + (size_t)c * kClumpFields;
            rec[kCId] = (double)vs[s0];
            rec[kCN] = (double)cnt;
            rec[kCMass] = m;
            rec[kCCom] = com.x;
            rec[kCCom + 1] = com.y;
            rec[kCCom + 2] = com.z;
            rec[kCVel] = cvel.x;
            rec[kCVel + 1] = cvel.y;
            rec[kCVel + 2] = cvel.z;
            rec[kCRho] = rho;
            rec[kCRadius] = radius;
            rec[kCKe] = ke;
            rec[kCPe] = 0.0;
            rec[kCL] = lx;
            rec[kCL + 1] = ly;
            rec[kCL + 2] = lz;
            rec[kCThermal] = fmax(ke - e_rot, 0.0);
            rec[kCBound] = 0.0;
            rec[kCOutside] = 0.0;
            rec[kCInst] = (double)minst[vs[s0]];
            const uint32_t nt = (cnt + 255u) / 256u;
            tcnt[c] = cnt >= 2 ? nt * (nt + 1u) / 2u : 0u;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) tcnt[nc] = 0u;
}

// Softened self potential energy, one 256 x 256 tile (rows a, columns b >= a) of one clump per block:
// sum over i < j of m_i m_j / sqrt(|x_j - x_i|^2 + s^2) (bhtree.rs:192-216, the exact branch).
__global__ __launch_bounds__(256) void k_clump_pe(uint32_t nc, const uint32_t* __restrict__ toff,
                                                  const uint32_t* __restrict__ cstart, const double4* __restrict__ cp,
                                                  const double* __restrict__ crec, const double* __restrict__ iparam,
                                                  double* __restrict__ tpe) {
#pragma clang fp contract(off)
    __shared__ double4 col[256];
    __shared__ double red[256];
    const uint32_t t = blockIdx.x;
    uint32_t lo = 0, hi = nc;  // largest c with toff[c] <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (toff[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t c = lo;
    const uint32_t s0 = cstart[c], cnt = cstart[c + 1] - s0, nt = (cnt + 255u) / 256u;
    uint32_t rem = t - toff[c], a = 0;
    while (rem >= nt - a) {
        rem -= nt - a;
        ++a;
    }
    const uint32_t b = a + rem;
    const double soft2 = iparam[(size_t)crec[(size_t)c * kClumpFields + kCInst] * kParamFields + kPSoft2];
    const uint32_t j = b * 256u + threadIdx.x;
    col[threadIdx.x] = j < cnt ? cp[s0 + j] : make_double4(0.0, 0.0, 0.0, 0.0);
    __syncthreads();
    const uint32_t i = a * 256u + threadIdx.x;
    double acc = 0.0;
    if (i < cnt) {
        const double4 P = cp[s0 + i];
        const uint32_t jn = min(256u, cnt - b * 256u);
        for (uint32_t jj = (a == b) ? threadIdx.x + 1u : 0u; jj < jn; ++jj) {
            const double4 Q = col[jj];
            const D3 d{Q.x - P.x, Q.y - P.y, Q.z - P.z};
            const double r2 = dot3(d, d) + soft2;
            acc += Q.w / sqrt(r2);
        }
        acc *= P.w;
    }
    const double s = block_sum(red, acc);
    if (threadIdx.x == 0) tpe[t] = s;
}

__global__ __launch_bounds__(256) void k_clump_finish(uint32_t nc, const uint32_t* __restrict__ toff,
                                                      const double* __restrict__ tpe, double* __restrict__ crec,
                                                      const double* __restrict__ iparam) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const uint32_t w0 = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const uint32_t ws = gridDim.x * (blockDim.x / 64);
    for (uint32_t c = w0; c < nc; c += ws) {
        double sum = 0.0;
        for (uint32_t t = toff[c] + lane; t < toff[c + 1]; t += 64) sum += tpe[t];
        sum = wave_sum_lane0(sum);
        if (lane == 0) {
            double* rec = crec + (size_t)c * kClumpFields;
            const double* pr = iparam + (size_t)rec[kCInst] * kParamFields;
            const double pe = rec[kCN] >= 2.0 ? -pr[kPG] * sum : 0.0;
            rec[kCPe] = pe;
            rec[kCBound] = (rec[kCKe] + pe < 0.0) ? 1.0 : 0.0;  // accretion.rs:279
            // accretion.rs:281-285, 304-306
            const double cr = pr[kPCRad];
            const double rho_c = pr[kPCMass] / (kFourThirdsPi * (cr * cr * cr));
            const double d_roche = 2.44 * cr * cbrt(rho_c / rec[kCRho]);
            const D3 d{rec[kCCom] - pr[kPCPos], rec[kCCom + 1] - pr[kPCPos + 1], rec[kCCom + 2] - pr[kPCPos + 2]};
            rec[kCOutside] = sqrt(dot3(d, d)) > d_roche ? 1.0 : 0.0;
        }
    }
}

__global__ __launch_bounds__(256) void k_clump_select(uint32_t nc, const double* __restrict__ crec, uint32_t inst,
                                                      uint32_t min_members, uint32_t* __restrict__ sel,
                                                      uint32_t* __restrict__ count) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double* rec = crec + (size_t)c * kClumpFields;
    if ((uint32_t)rec[kCInst] != inst || rec[kCN] < (double)min_members) return;
    sel[atomicAdd(count, 1u)] = c;
}

__global__ __launch_bounds__(256) void k_clump_gather(uint32_t count, const uint32_t* __restrict__ sel,
                                                      const double* __restrict__ crec, double* __restrict__ srec) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count * (uint32_t)kClumpFields) return;
    const uint32_t k = e / kClumpFields, f = e % kClumpFields;
    srec[e] = crec[(size_t)sel[k] * kClumpFields + f];
}

// impact-run's clump counts (main.rs:838-866): clumps of >= 2 members, the self-bound ones, and the mass of
// the largest that accretes (positive f64 bits order as integers).
// Clumps arrive in instance order (a clump's id is its first member slot and member slots are grouped by
// instance), so a wave's clumps nearly always share one instance: each wave folds the counts and the
// largest mass of every instance it holds and one lane per instance updates the counters.
__global__ __launch_bounds__(256) void k_clump_summary(uint32_t nc, const double* __restrict__ crec,
                                                       unsigned long long* __restrict__ csum) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    uint32_t inst = 0xffffffffu;
    bool bound = false;
    unsigned long long mbits = 0ull;
    if (c < nc) {
        const double* rec = crec + (size_t)c * kClumpFields;
        if (rec[kCN] >= 2.0) {
            inst = (uint32_t)rec[kCInst];
            bound = rec[kCBound] != 0.0;
            if (bound && rec[kCOutside] != 0.0) mbits = (unsigned long long)__double_as_longlong(rec[kCMass]);
        }
    }
    unsigned long long pending = __ballot(inst != 0xffffffffu);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t li = __shfl(inst, leader);
        const bool mine = inst == li;
        const unsigned long long mm = __ballot(mine), bb = __ballot(mine && bound);
        unsigned long long v = mine ? mbits : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long t = __shfl_xor(v, o);
            v = t > v ? t : v;
        }
        if (lane == leader) {
            unsigned long long* cs = csum + 3 * (size_t)li;
            atomicAdd(cs, (unsigned long long)__popcll(mm));
            if (bb) atomicAdd(cs + 1, (unsigned long long)__popcll(bb));
            if (v) atomicMax(cs + 2, v);
        }
        pending &= ~mm;
    }
}

__global__ void k_clear_u64(unsigned long long* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

__global__ void k_clear_words(uint32_t* word) {
    if (threadIdx.x < 8) word[threadIdx.x] = 0u;
}

}  // namespace

void launch_meas_live(const Dev& d, Meas& m, uint32_t ninst, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, s, m.word);
    hipLaunchKernelGGL(k_meas_live, dim3(m.nchunks), dim3(kMT), 0, s, m.chunks, d.posh, d.velu, d.mass, m.part);
    hipLaunchKernelGGL(k_meas_view1, dim3(ninst), dim3(kMT), 0, s, m.ichunk, m.part, m.view, m.cross);
}

void launch_meas_remnant(const Dev& d, Meas& m, uint32_t ninst, hipStream_t s) {
    const uint32_t n = d.n;
    const uint32_t* inst = ninst > 1 ? d.inst : nullptr;
    const uint32_t* ioff = ninst > 1 ? d.ioff : nullptr;
    hipLaunchKernelGGL(k_meas_rkey, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, inst, d.posh, d.velu, d.mass, m.view,
                       m.k0, m.khi, m.v0);
    // (lo word, then hi word) LSD passes: the order of the f64 radius; then stable by instance
    for (int sh = 0; sh < 32; sh += 16) {
        launch_radix_pass_n(n, d.hist, m.k0, m.v0, m.k1, m.v1, sh, s);
        launch_radix_pass_n(n, d.hist, m.k1, m.v1, m.k0, m.v0, sh + 8, s);
    }
    hipLaunchKernelGGL(k_gather_u32, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, m.khi, m.v0, m.k1);
    for (int sh = 0; sh < 32; sh += 16) {
        launch_radix_pass_n(n, d.hist, m.k1, m.v0, m.k0, m.v1, sh, s);
        launch_radix_pass_n(n, d.hist, m.k0, m.v1, m.k1, m.v0, sh + 8, s);
    }
    uint32_t* vs = m.v0;
    if (inst) {
        hipLaunchKernelGGL(k_gather_u32, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, inst, m.v0, m.k0);
        launch_radix_pass_n(n, d.hist, m.k0, m.v0, m.k1, m.v1, 0, s);  // at most 256 instances
        vs = m.v1;
    }
    hipLaunchKernelGGL(k_meas_scan, dim3(m.nchunks), dim3(kMT), 0, s, m.chunks, m.cinst, vs, d.posh, d.velu, d.mass,
                       m.view, m.scum, m.coff);
    hipLaunchKernelGGL(k_meas_coff, dim3(ninst), dim3(kMT), 0, s, m.ichunk, m.coff);
    hipLaunchKernelGGL(k_meas_cross, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, inst, ioff, m.ichunk, vs, d.posh,
                       d.velu, d.mass, m.view, m.scum, m.coff, m.cross);
    hipLaunchKernelGGL(k_meas_view2, dim3(cdiv_m(ninst, 64)), dim3(64), 0, s, ninst, ioff, m.ichunk, vs, d.posh,
                       d.velu, d.mass, m.scum, m.coff, m.cross, m.view);
}

void launch_meas_classify(const Dev& d, Meas& m, uint32_t ninst, double clump_soft, double link_factor,
                          hipStream_t s) {
    const uint32_t n = d.n;
    const uint32_t* inst = ninst > 1 ? d.inst : nullptr;
    const uint32_t* ioff = ninst > 1 ? d.ioff : nullptr;
    hipLaunchKernelGGL(k_meas_classify, dim3(m.nchunks), dim3(kMT), 0, s, m.chunks, m.cinst, d.posh, d.velu, d.mass,
                       d.prov, m.view, m.part, m.slot);
    hipLaunchKernelGGL(k_meas_view3, dim3(ninst), dim3(kMT), 0, s, m.ichunk, m.part, m.view, m.iparam, clump_soft,
                       link_factor);
    hipLaunchKernelGGL(k_meas_mbase, dim3(1), dim3(kMbT), 0, s, ninst, m.nchunks, m.ichunk, m.part, m.mbase, m.moff,
                       m.word);
    hipLaunchKernelGGL(k_meas_gather, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, inst, ioff, m.ichunk, d.posh, d.velu,
                       d.mass, d.rho, m.slot, m.mbase, m.mp, m.mv, m.minst, m.midx);
}

void launch_fof(Meas& m, uint32_t n, uint32_t* hist, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, s, m.word + 8);
    hipLaunchKernelGGL(k_fof_key, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, m.mp, m.minst, m.iparam, m.k0, m.v0,
                       m.parent);
    for (int sh = 0; sh < 32; sh += 16) {
        launch_radix_pass_n(n, hist, m.k0, m.v0, m.k1, m.v1, sh, s);
        launch_radix_pass_n(n, hist, m.k1, m.v1, m.k0, m.v0, sh + 8, s);
    }
    hipLaunchKernelGGL(k_fof_link, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, m.k0, m.v0, m.mp, m.minst, m.iparam,
                       m.parent, m.word + 9);
    hipLaunchKernelGGL(k_fof_label, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, m.parent, m.label, m.k0, m.v0);
    // stable sort by label (the label is a member slot < n): members of a clump stay in slot order
    uint32_t *ki = m.k0, *vi = m.v0, *ko = m.k1, *vo = m.v1;
    for (int sh = 0; sh < 32 && (sh == 0 || ((n - 1) >> sh) != 0); sh += 8) {
        launch_radix_pass_n(n, hist, ki, vi, ko, vo, sh, s);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    m.ks = ki;
    m.vs = vi;
    hipLaunchKernelGGL(k_clump_flag, dim3(cdiv_m(n + 1, 256)), dim3(256), 0, s, n, ki, m.cflag);
    launch_exclusive_scan(m.cflag, n + 1, s);
    hipLaunchKernelGGL(k_clump_emit, dim3(cdiv_m(n, 256)), dim3(256), 0, s, n, ki, vi, m.cflag, m.cstart, m.mp, m.mv,
                       m.cp, m.cv, m.word);
}

void launch_clump_stats(Meas& m, uint32_t n_clumps, hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>(cdiv_m(n_clumps, 4), 8192));
    hipLaunchKernelGGL(k_clump_stats, dim3(blocks), dim3(256), 0, s, n_clumps, m.cstart, m.vs, m.minst, m.cp, m.cv,
                       m.crec, m.tcnt);
    launch_exclusive_scan(m.tcnt, n_clumps + 1, s);
}

void launch_clump_pe(Meas& m, uint32_t n_clumps, uint32_t n_tiles, hipStream_t s) {
    if (n_tiles > 0)
        hipLaunchKernelGGL(k_clump_pe, dim3(n_tiles), dim3(256), 0, s, n_clumps, m.tcnt, m.cstart, m.cp, m.crec,
                           m.iparam, m.tpe);
    const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>(cdiv_m(n_clumps, 4), 8192));
    hipLaunchKernelGGL(k_clump_finish, dim3(blocks), dim3(256), 0, s, n_clumps, m.tcnt, m.tpe, m.crec, m.iparam);
}

void launch_clump_select(Meas& m, uint32_t n_clumps, uint32_t inst, uint32_t min_members, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, s, m.word + 16);
    if (n_clumps > 0)
        hipLaunchKernelGGL(k_clump_select, dim3(cdiv_m(n_clumps, 256)), dim3(256), 0, s, n_clumps, m.crec, inst,
                           min_members, m.sel, m.word + 16);
}

void launch_clump_summary(Meas& m, uint32_t n_clumps, uint32_t ninst, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(256), 0, s, m.csum, 3u * ninst);
    if (n_clumps > 0)
        hipLaunchKernelGGL(k_clump_summary, dim3(cdiv_m(n_clumps, 256)), dim3(256), 0, s, n_clumps, m.crec, m.csum);
}

void launch_clump_gather(Meas& m, uint32_t count, hipStream_t s) {
    if (count > 0)
        hipLaunchKernelGGL(k_clump_gather, dim3(cdiv_m(count * (uint32_t)kClumpFields, 256)), dim3(256), 0, s, count,
                           m.sel, m.crec, m.srec);
}

}  // namespace isph
