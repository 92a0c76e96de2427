// SPH density, per-particle FP64 Tillotson EOS, Courant signal, pressure + artificial-viscosity force with
// du/dt, the KDK integrators, AoS pack/unpack and the f64 diagnostics.
//
// Reference: sph_step.wgsl:181-216 (density), :219-296 (forces + du/dt), :304-333 (KDK), :364-371 (signal);
// the CPU path hydrostatic.rs:338-480 (same physics in f64, EOS once per particle with the eos.rs:88-96
// sound speed). Neighbours are found through the LBVH instead of the reference's 2^15-cell hash table: one
// wave walks the tree for 64 Morton-consecutive particles, keeping every node whose box comes within
// (h_group_max + h_node_max)/2 of the group box, so the candidate set contains every pair with
// r < (h_i + h_j)/2 exactly once; each lane then applies the exact radius test, as the reference's
// cell-membership-guarded 27-cell scan does. Waves take walk groups (compact subtrees of <= 64 particles)
// from a work queue. Pressure and sound speed are computed once per particle (the
// reference shader re-evaluates six Tillotson calls per neighbour pair, sph_step.wgsl:273-275).
#include "walk.h"

namespace isph {


// AoS record, identical layout to isph_particle (gpu_sph.rs:27-37).
struct ParticleRec {
    float px, py, pz, h;
    float vx, vy, vz, u;
    float mass;
    uint32_t mat;
    float rho;
    uint32_t prov;
};
static_assert(sizeof(ParticleRec) == 48, "SphParticle layout");

// Per-particle EOS: pressure / rho^2 and sound speed from the f64 Tillotson (eos.rs:88-96); returns the
// particle's Courant signal h / max(c + |v|, 1).
__device__ __forceinline__ float eos_point(uint32_t t, double rho, float h, float4 v, uint32_t m,
                                           const Till* __restrict__ eos, float* spr, float* scs, double* spr64,
                                           double* scs64) {
    const Till& T = eos[m];
    double p, c2;
    till_p_c2(T, rho, (double)v.w, p, c2);
    double c = sqrt(c2);
    spr[t] = (float)(p / (rho * rho));
    scs[t] = (float)c;
    if (spr64) {
        spr64[t] = p / (rho * rho);
        scs64[t] = c;
    }
    float speed = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return h / fmaxf((float)c + speed, 1.0f);  // sph_step.wgsl:370, hydrostatic.rs:477
}

// Candidate-list cache: the density walk stores, per group, the candidates that pair with some member (in
// walk order, packed into chunks of 64 slots) and each lane's pair mask over a chunk; the force pass replays
// them.
struct CandCache {
    uint32_t* head;
    uint32_t* next;
    uint32_t* count;
    uint32_t* pool;
    uint64_t* mask;
    uint32_t blocks;
};

// Wave-uniform writer state: the current block of chunks and the group's first and last chunk. A wave
// starts in block `blockIdx.x` (the first kWalkWaves blocks are one per wave) and takes further blocks
// with one atomic per kChunksPerBlock chunks. A group whose chunks run out is marked for the walk.
struct CacheWriter {
    uint32_t blk = blockIdx.x, used = 0, first = kNoChunk, last = kNoChunk;
    bool ok = true;
    __device__ __forceinline__ void begin() {
        first = last = kNoChunk;
        ok = true;
    }
    __device__ __forceinline__ void put(const CandCache& cc, uint32_t* scal, const uint32_t* cand, int cnt,
                                        int lane) {
        if (!ok) return;
        if (blk >= cc.blocks) {  // only below kWalkWaves blocks: a capped cache (isph_set_candidate_cache)
            ok = false;
            return;
        }
        if (used == kChunksPerBlock) {
            uint32_t b = 0;
            if (lane == 0) b = kWalkWaves + atomicAdd(&scal[kPoolHead], 1u);
            b = __builtin_amdgcn_readfirstlane(b);
            if (b >= cc.blocks) {
                ok = false;
                return;
            }
            blk = b;
            used = 0;
        }
        const uint32_t id = blk * kChunksPerBlock + used++;
        if (lane < cnt) cc.pool[(size_t)id * kWave + lane] = cand[lane];
        if (lane == 0) {
            cc.count[id] = (uint32_t)cnt;
            if (last == kNoChunk) first = id;
            else cc.next[last] = id;
        }
        last = id;
    }
    __device__ __forceinline__ void end(const CandCache& cc, uint32_t gi, int lane) {
        if (lane != 0) return;
        if (!ok) {
            cc.head[gi] = kChunkOverflow;
            return;
        }
        if (last != kNoChunk) cc.next[last] = kNoChunk;
        cc.head[gi] = first;
    }
};

// OR over the 64 lanes (all active), returned wave-uniform: DPP row shifts within each row of 16, then
// row_bcast:15 and row_bcast:31 carry rows 0 | 1 and 2 | 3 into lane 63.
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ uint64_t wave_or_u64(uint64_t v) {
    return ((uint64_t)wave_or_u32((uint32_t)(v >> 32)) << 32) | wave_or_u32((uint32_t)v);
}
// The bits of x selected by m, packed towards bit 0 in order (bit j of x & m lands at the count of m's bits
// below j): the six-step parallel compress (Warren, Hacker's Delight, 2nd ed., section 7-4). Everything that
// depends on m alone is wave-uniform when m is.
__device__ __forceinline__ uint64_t compress_bits(uint64_t x, uint64_t m) {
    x &= m;
    uint64_t mk = ~m << 1;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        uint64_t mp = mk ^ (mk << 1);
        mp ^= mp << 2;
        mp ^= mp << 4;
        mp ^= mp << 8;
        mp ^= mp << 16;
        mp ^= mp << 32;
        const uint64_t mv = mp & m;
        m = (m ^ mv) | (mv >> (1 << i));
        const uint64_t t = x & mv;
        x = (x ^ t) | (t >> (1 << i));
        mk &= ~mp;
    }
    return x;
}

// REF (the f64 reference form, F64 too): the pair terms take f64 coordinates, each particle's f32 copy plus
// its f32 residual against the f64 state (x64 = x32 + lo), so differences are the f64 state's to ~1e-16.
template <bool F64, bool REF = false>
__global__ __launch_bounds__(64, F64 ? 6 : 8) void k_density(const float4* __restrict__ spos,
                                                             const float* __restrict__ sh, uint32_t n, uint32_t K, uint32_t L, const SphNode* __restrict__ nS,
                                                const uint32_t* __restrict__ order,
                                                float* __restrict__ srho, float* __restrict__ rho_orig,
                                                double* __restrict__ srho64, uint32_t* scal,
                                                unsigned long long* counters, const uint2* __restrict__ groups,
                                                const uint2* __restrict__ nrange, uint32_t expand,
                                                const uint32_t* __restrict__ sinst,
                                                const uint32_t* __restrict__ iroot, CandCache cc,
                                                const float4* __restrict__ slo) {
    __shared__ uint32_t stk[kSphStack];
    __shared__ uint32_t cand[kSphBuf];
    __shared__ float4 cP[kWave];  // x, y, z, h
    __shared__ float cM[kWave];   // mass
    __shared__ float4 cL[REF ? kWave : 1];  // position residuals (reference form)
    __shared__ uint32_t own[kWave];
    __shared__ uint32_t pend[2 * kWave];  // candidate ids of the chunk being filled, and the overflow
    const int lane = threadIdx.x;
    own[lane] = kWave;
    unsigned long long tested_all = 0;
    CacheWriter cw;
    WalkQueue queue(scal + kQueueDensity, scal[kGroupsSph], 1);
    uint32_t gi;
    while (queue.next(gi)) {
    const uint2 grp = groups[gi];
    // a group of at most 32 particles runs on both wave halves (pair_mask_half), member = lane mod 32
    const bool split = grp.y <= (uint32_t)(kWave / 2);
    const int member = split ? (lane & (kWave / 2 - 1)) : lane;
    const uint32_t t = grp.x + (uint32_t)member;
    const bool valid = (uint32_t)member < grp.y;
    const uint32_t tc = valid ? t : grp.x;
    __syncthreads();
    const float4 me = spos[tc];
    const float4 melo = REF ? slo[tc] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float hi = sh[tc];
    const float qi = 0.5f * hi;  // half smoothing length, as the candidates carry it
    const float gx0 = uniform_f(wave_min(me.x)), gy0 = uniform_f(wave_min(me.y)), gz0 = uniform_f(wave_min(me.z));
    const float gx1 = uniform_f(wave_max(me.x)), gy1 = uniform_f(wave_max(me.y)), gz1 = uniform_f(wave_max(me.z));
    const float hg = uniform_f(wave_max(hi));
    const float qg = 0.5f * hg;
    // Plain: f32 pair arithmetic and sum (the reference shader's arithmetic). Mixed (F64): the density feeds
    // the Tillotson pressure, whose cold branch P = A (rho/rho0 - 1) + ... cancels near rho0 and amplifies a
    // relative density error by A / P (~1e3), so the true-pair kernel sums run in f64, from f64 differences
    // of the f32 coordinates (only ~50 pairs per particle pass the radius test).
    float rho = 0.f;
    double rho64 = 0.0;
    unsigned long long tested = 0;
    // The candidate cache keeps only the candidates that pair with some member of the group, in walk order,
    // packed into chunks of 64 slots (`pc` slots of the pending chunk are filled); `pm` is this lane's pair
    // mask over those slots. A lane's hits keep their order, so the force pass's sums are the walk's.
    uint64_t pm = 0ull;
    int pc = 0;
    cw.begin();
    neighbour_walk(stk, cand, n, K, L, nS, expand, gx0, gy0, gz0, gx1, gy1, gz1, hg,
                   counters ? counters + (size_t)blockIdx.x * kCounterSlots + 8 : nullptr, [&](int cnt) {
        // 0) Candidates that pass the prefilter (near_group), packed in walk order (cand[] in place: every
        //    lane has read its entry before any is overwritten).
        bool pass = false;
        uint32_t j = 0;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        float qj = 0.f;
        if (lane < cnt) {
            j = cand[lane];
            p = spos[j];
            qj = 0.5f * sh[j];
            pass = near_group(p, qj, gx0, gy0, gz0, gx1, gy1, gz1, qg);
        }
        const uint64_t sv = __ballot(pass);
        const int nc = __popcll(sv);
        if (pass) {
            const int at = prefix_count(sv);
            cand[at] = j;
            cP[at] = make_float4(p.x, p.y, p.z, qj);
            cM[at] = p.w;
            if (REF) cL[at] = slo[j];
            note_owner(own, j, grp.x, grp.y, at);
        }
        __syncthreads();
        // 1) The radius test against every candidate (LDS broadcast reads): one bit per true pair, the
        //    lane's own particle excluded (sph_step.wgsl:210).
        const uint64_t hit = (split ? pair_mask_half(cP, nc, me, qi, lane >= kWave / 2) : pair_mask(cP, nc, me, qi)) &
                             ~self_bit(own, cand, nc, member, t);
        // The kept candidates (a pair of some member) go to the pending chunk in walk order.
        const uint64_t keep = cw.ok ? wave_or_u64(valid ? hit : 0ull) : 0ull;
        if (lane < nc && ((keep >> lane) & 1ull)) pend[pc + prefix_count(keep)] = cand[lane];
        // 2) The kernel sum over this lane's hits only, in candidate order. Lanes walk their own hits in
        //    step, so the wave runs as many iterations as its busiest lane instead of one per candidate with
        //    most lanes idle.
        float acc = 0.f;
        uint64_t rest = valid ? hit : 0ull;
        while (__ballot(rest != 0ull)) {
            if (rest) {
                const int jj = __builtin_ctzll(rest);
                rest &= rest - 1ull;
                const float4 pj = cP[jj];
                const float mj = cM[jj];
                if (F64) {
                    double ddx = (double)me.x - (double)pj.x, ddy = (double)me.y - (double)pj.y,
                           ddz = (double)me.z - (double)pj.z;
                    if (REF) {
                        const float4 lj = cL[jj];
                        ddx += (double)melo.x - (double)lj.x;
                        ddy += (double)melo.y - (double)lj.y;
                        ddz += (double)melo.z - (double)lj.z;
                    }
                    double inv_h = nr_rcp((double)qi + (double)pj.w);
                    const double rr = ddx * ddx + ddy * ddy + ddz * ddz;
                    double q = (rr > 0.0 ? rr * nr_rsq(rr) : 0.0) * inv_h;
                    double sig = (8.0 / 3.14159265358979323846) * (inv_h * inv_h * inv_h);
                    double w = q < 0.5 ? sig * (6.0 * (q * q * q - q * q) + 1.0)
                                       : (q < 1.0 ? sig * 2.0 * (1.0 - q) * (1.0 - q) * (1.0 - q) : 0.0);
                    rho64 += (double)mj * w;
                } else {
                    float dx = me.x - pj.x, dy = me.y - pj.y, dz = me.z - pj.z;
                    float r2 = dx * dx + dy * dy + dz * dz;
                    float hij = qi + pj.w;
                    float r = hw_sqrt(r2);
                    float inv_h = hw_rcp(hij);
                    acc += mj * spline_w(r * inv_h, inv_h);
                }
            }
        }
        // the force pass of this evaluation replays the kept candidates with these masks (its pair set is
        // the same)
        if (cw.ok) {
            const int k = __popcll(keep);
            const uint64_t cm = compress_bits(valid ? hit : 0ull, keep);  // the hits' slots among the kept
            pm |= cm << pc;  // pc < 64
            const uint64_t carry = pc + k > kWave ? cm >> (kWave - pc) : 0ull;  // then pc >= 1
            pc += k;
            if (pc >= kWave) {
                __syncthreads();
                cw.put(cc, scal, pend, kWave, lane);
                if (cw.ok) cc.mask[(size_t)cw.last * kWave + lane] = pm;
                pm = carry;
                pc -= kWave;
                __syncthreads();
                if (lane < pc) pend[lane] = pend[lane + kWave];
            }
        }
        rho += acc;
        tested += nc;
    }, sinst ? iroot[sinst[grp.x]] : 0u);
    if (cw.ok && pc > 0) {
        __syncthreads();
        cw.put(cc, scal, pend, pc, lane);
        if (cw.ok) cc.mask[(size_t)cw.last * kWave + lane] = pm;
    }
    cw.end(cc, gi, lane);
    if (split) {  // the upper half's sum joins its member's lane
        if (F64) {
            rho += __shfl_down(rho, kWave / 2);
            rho64 += __shfl_down(rho64, kWave / 2);
        } else {
            // Through LDS (the candidate masses are free here): a cross-lane shuffle's lane address would be
            // held across the walk and, at the 64-VGPR budget of 8 waves per SIMD, spilled to scratch.
            __syncthreads();
            if (lane >= kWave / 2) cM[lane - kWave / 2] = rho;
            __syncthreads();
            if (lane < kWave / 2) rho += cM[lane];
        }
    }
    const bool owner = valid && (!split || lane < kWave / 2);
    const bool live = owner && me.w > 0.f;  // retired (mass 0) particles keep their density (:185-189)
    double rho_i;
    if (live) {
        if (F64) {
            rho_i = (double)me.w * (8.0 / (3.14159265358979323846 * (double)hi * (double)hi * (double)hi)) + rho64;
        } else {
            float inv_h = 1.0f / hi;
            rho_i = (double)(me.w * spline_w(0.f, inv_h) + rho);
        }
        rho_orig[order[t]] = (float)rho_i;
    } else {
        rho_i = owner ? (double)rho_orig[order[t]] : 1.0;
    }
    // The per-particle EOS runs in its own pass (k_eos): its f64 Tillotson evaluations would otherwise set
    // this walk's register budget and halve the resident waves.
    if (owner) {
        srho[t] = (float)rho_i;
        if (F64) srho64[t] = rho_i;
    }
    tested_all += tested * grp.y;
    }
    if (lane == 0 && counters) counters[(size_t)blockIdx.x * kCounterSlots + 3] += tested_all;
}

// Per-particle EOS and the Courant signal minimum, after the density walk (or on the reused density when
// positions have not moved). Grid-stride with a block-level minimum: one atomic per block, since
// same-address atomics from every wave would serialise.
constexpr int kEosThreads = 256;
constexpr uint32_t kEosBlocks = 1024;
__global__ __launch_bounds__(kEosThreads) void k_eos(const float* __restrict__ sh, const float4* __restrict__ svel,
                                                     const uint32_t* __restrict__ smat, uint32_t n,
                                                     const Till* __restrict__ eos, const float* __restrict__ srho,
                                                     float* __restrict__ spr, float* __restrict__ scs,
                                                     const double* __restrict__ srho64, double* __restrict__ spr64,
                                                     double* __restrict__ scs64, uint32_t* scal,
                                                     const uint32_t* __restrict__ sinst, uint32_t* isig) {
    __shared__ float red[kEosThreads / kWave];
    float sig = INFINITY;
    if (sinst) {  // ensemble: a Courant minimum per instance (waves inside one instance reduce first)
        for (uint32_t base = blockIdx.x * kEosThreads; base < n; base += gridDim.x * kEosThreads) {
            const uint32_t t = base + threadIdx.x;
            const bool valid = t < n;
            float v = INFINITY;
            if (valid) {
                const double rho = srho64 ? srho64[t] : (double)srho[t];
                v = eos_point(t, rho, sh[t], svel[t], smat[t], eos, spr, scs, spr64, scs64);
            }
            const uint32_t k = valid ? sinst[t] : 0xffffffffu;
            const uint32_t k0 = __builtin_amdgcn_readfirstlane(k);
            if (__ballot(valid && k != k0) == 0ull) {
                v = wave_min(v);
                if (lane_id() == 0 && k0 != 0xffffffffu && v < INFINITY) atomicMin(&isig[k0], __float_as_uint(v));
            } else if (valid && v < INFINITY) {
                atomicMin(&isig[k], __float_as_uint(v));
            }
        }
        return;
    }
    for (uint32_t t = blockIdx.x * kEosThreads + threadIdx.x; t < n; t += gridDim.x * kEosThreads) {
        const double rho = srho64 ? srho64[t] : (double)srho[t];
        sig = fminf(sig, eos_point(t, rho, sh[t], svel[t], smat[t], eos, spr, scs, spr64, scs64));
    }
    sig = wave_min(sig);
    if (lane_id() == 0) red[threadIdx.x / kWave] = sig;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kEosThreads / kWave; ++w) sig = fminf(sig, red[w]);
        if (sig < INFINITY) atomicMin(&scal[6], __float_as_uint(sig));
    }
}

// Resident waves per SIMD: five for the plain form, four for the f64 form (up to 128 VGPRs, no spilling).
template <bool F64, bool REF = false>
__global__ __launch_bounds__(64, F64 ? 4 : 5) void k_force(const float4* __restrict__ spos, const float* __restrict__ sh,
                                              const float4* __restrict__ svel, const float* __restrict__ srho,
                                              const float* __restrict__ spr, const float* __restrict__ scs, uint32_t n,
                                              uint32_t K, uint32_t L, const SphNode* __restrict__ nS,
                                              const uint32_t* __restrict__ order,
                                              const float4* __restrict__ sgrav, float4* __restrict__ acc_orig,
                                              const double* __restrict__ srho64, const double* __restrict__ spr64,
                                              const double* __restrict__ scs64, float av_alpha, float av_beta, unsigned long long* counters,
                                              const uint2* __restrict__ groups, uint32_t* scal,
                                              const uint2* __restrict__ nrange, uint32_t expand,
                                              double* __restrict__ acc64, const uint32_t* __restrict__ sinst,
                                              const uint32_t* __restrict__ iroot, Boundary bnd,
                                              const uint32_t* __restrict__ chead, const uint32_t* __restrict__ cnext,
                                              const uint32_t* __restrict__ ccnt, const uint32_t* __restrict__ cpool,
                                              const uint64_t* __restrict__ cmask, const float4* __restrict__ slo,
                                              const float4* __restrict__ svlo, const double4* __restrict__ sgrav64) {
    __shared__ uint32_t stk[kSphStack];
    __shared__ uint32_t cand[kSphBuf];
    __shared__ float4 cP[kWave];  // x, y, z, h
    __shared__ float4 cV[kWave];  // vx, vy, vz, rho
    __shared__ float4 cS[kWave];  // mass, P/rho^2, c
    __shared__ double cR64[F64 ? kWave : 1];
    __shared__ double cPr64[F64 ? kWave : 1];
    __shared__ double cC64[F64 ? kWave : 1];
    __shared__ float4 cL[REF ? kWave : 1];   // position residuals (reference form)
    __shared__ float4 cVL[REF ? kWave : 1];  // velocity residuals
    __shared__ uint32_t own[kWave];
    const int lane = threadIdx.x;
    own[lane] = kWave;
    unsigned long long tested_all = 0;
    WalkQueue queue(scal + kQueueForce, scal[kGroupsSph]);
    uint32_t gi;
    while (queue.next(gi)) {
    const uint2 grp = groups[gi];
    // as in the density pass: a group of at most 32 particles runs on both wave halves, member = lane mod 32
    const bool split = grp.y <= (uint32_t)(kWave / 2);
    const int member = split ? (lane & (kWave / 2 - 1)) : lane;
    const uint32_t t = grp.x + (uint32_t)member;
    const bool valid = (uint32_t)member < grp.y;
    const uint32_t tc = valid ? t : grp.x;
    __syncthreads();
    const float4 me = spos[tc];
    const float hi = sh[tc];
    const float qi = 0.5f * hi;  // half smoothing length, as the candidates carry it
    const float4 vi = svel[tc];
    const float4 melo = REF ? slo[tc] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vilo = REF ? svlo[tc] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float rhoi = srho[tc], pri = spr[tc], ci = scs[tc];
    const float gx0 = uniform_f(wave_min(me.x)), gy0 = uniform_f(wave_min(me.y)), gz0 = uniform_f(wave_min(me.z));
    const float gx1 = uniform_f(wave_max(me.x)), gy1 = uniform_f(wave_max(me.y)), gz1 = uniform_f(wave_max(me.z));
    const float hg = uniform_f(wave_max(hi));
    const float qg = 0.5f * hg;
    // Plain: f32 pair arithmetic and sums (the reference shader's). Mixed (F64): the true pairs (about 50
    // per particle) are evaluated and summed in f64 from f64 coordinate differences, with the f64
    // density and P / rho^2 of the EOS pass; du/dt cancels across pairs, so f32 pair terms cap its
    // relative accuracy near 1e-4.
    float ax = 0.f, ay = 0.f, az = 0.f, de = 0.f;
    double ax64 = 0.0, ay64 = 0.0, az64 = 0.0, de64 = 0.0;
    const double rhoi64 = F64 ? srho64[tc] : 0.0, pri64 = F64 ? spr64[tc] : 0.0, ci64 = F64 ? scs64[tc] : 0.0;
    unsigned long long tested = 0;
    // One candidate's data, loaded into registers (`fetch`) and staged in LDS (`put`).
    struct Cand {
        float4 p, v;
        float h, pr, cs;
    };
    auto fetch = [&](uint32_t j) {
        Cand c;
        c.p = spos[j];
        c.h = sh[j];
        c.v = svel[j];
        c.v.w = srho[j];
        c.pr = spr[j];
        c.cs = scs[j];
        return c;
    };
    auto put = [&](const Cand& c, uint32_t j, int at) {
        cP[at] = make_float4(c.p.x, c.p.y, c.p.z, 0.5f * c.h);
        cV[at] = c.v;
        cS[at] = make_float4(c.p.w, c.pr, c.cs, 0.f);
        if (F64) {
            cR64[at] = srho64[j];
            cPr64[at] = spr64[j];
            cC64[at] = scs64[j];
        }
        if (REF) {
            cL[at] = slo[j];
            cVL[at] = svlo[j];
        }
    };
    // The pair terms of one staged batch over `rest`, this lane's hits. 1) The radius test is r < (h_i +
    // h_j) / 2 in f32 (sph_step.wgsl:262-272), the density's own test, so a replayed chunk takes the
    // density's mask. 2) Each lane walks its own hits in candidate order; coincident pairs (r^2 <= 1e-18,
    // sph_step.wgsl:268) are skipped there.
    auto pairs = [&](int cnt, uint64_t rest) {
        while (__ballot(rest != 0ull)) {
            if (!rest) continue;
            const int jj = __builtin_ctzll(rest);
            rest &= rest - 1ull;
            const float4 pj = cP[jj];
            const float4 sj = cS[jj];
            if (F64) {
                double dx = (double)me.x - (double)pj.x, dy = (double)me.y - (double)pj.y,
                       dz = (double)me.z - (double)pj.z;
                if (REF) {
                    const float4 lj = cL[jj];
                    dx += (double)melo.x - (double)lj.x;
                    dy += (double)melo.y - (double)lj.y;
                    dz += (double)melo.z - (double)lj.z;
                }
                double r2 = dx * dx + dy * dy + dz * dz;
                double hij = (double)qi + (double)pj.w;
                if (r2 > 1.0e-18) {
                    float4 vj = cV[jj];
                    double inv_r = nr_rsq(r2);
                    double inv_h = nr_rcp(hij);
                    double q = r2 * inv_r * inv_h;
                    double dvx = (double)vi.x - (double)vj.x, dvy = (double)vi.y - (double)vj.y,
                           dvz = (double)vi.z - (double)vj.z;
                    if (REF) {
                        const float4 wj = cVL[jj];
                        dvx += (double)vilo.x - (double)wj.x;
                        dvy += (double)vilo.y - (double)wj.y;
                        dvz += (double)vilo.z - (double)wj.z;
                    }
                    double vr = dvx * dx + dvy * dy + dvz * dz;
                    double pij = 0.0;
                    if (vr < 0.0) {
                        // Monaghan viscosity with mu = hij vr / D, D = r^2 + 0.01 hij^2, and one reciprocal
                        // for both divisions: pij = 2 hij vr (beta hij vr - alpha cbar D) / (D^2 (rho_i + rho_j))
                        const double D = r2 + 0.01 * hij * hij;
                        const double cbar = 0.5 * (ci64 + cC64[jj]);
                        const double hv = hij * vr;
                        pij = 2.0 * hv * ((double)av_beta * hv - (double)av_alpha * cbar * D) *
                              nr_rcp(D * D * (rhoi64 + cR64[jj]));
                    }
                    double coeff = pri64 + cPr64[jj] + pij;
                    double ih2 = inv_h * inv_h;
                    double sig = (8.0 / 3.14159265358979323846) * (ih2 * ih2);
                    double dw = q < 0.5 ? sig * (18.0 * q * q - 12.0 * q) : sig * -6.0 * (1.0 - q) * (1.0 - q);
                    // grad W = dpos dw / r, so dv . grad W = vr dw / r: one scalar weight per pair
                    const double f = coeff * (double)sj.x * (dw * inv_r);
                    ax64 = fma(-dx, f, ax64);
                    ay64 = fma(-dy, f, ay64);
                    az64 = fma(-dz, f, az64);
                    de64 = fma(0.5 * f, vr, de64);
                }
            } else {
                float dx = me.x - pj.x, dy = me.y - pj.y, dz = me.z - pj.z;  // dpos = pos_i - pos_j
                float r2 = dx * dx + dy * dy + dz * dz;
                if (!(r2 > 1.0e-18f)) continue;
                float hij = qi + pj.w;
                float4 vj = cV[jj];
                float r = hw_sqrt(r2);
                float inv_h = hw_rcp(hij);
                float dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
                float vr = dvx * dx + dvy * dy + dvz * dz;
                float pij = 0.f;
                if (vr < 0.f) {  // Monaghan viscosity (:279-284)
                    float mu = hij * vr * hw_rcp(r2 + 0.01f * hij * hij);
                    float cbar = 0.5f * (ci + sj.z);
                    float rbar = 0.5f * (rhoi + vj.w);
                    pij = (-av_alpha * cbar * mu + av_beta * mu * mu) * hw_rcp(rbar);
                }
                float coeff = pri + sj.y + pij;
                float gs = spline_dw(r * inv_h, inv_h) * hw_rcp(r);
                float gx = dx * gs, gy = dy * gs, gz = dz * gs;
                float f = -coeff * sj.x;
                ax += gx * f;
                ay += gy * f;
                az += gz * f;
                de += 0.5f * sj.x * coeff * (dvx * gx + dvy * gy + dvz * gz);
            }
        }
        tested += cnt;
    };
    auto flush = [&](int cnt) {  // a batch of the walk: load, prefilter and pack as the density walk, test
        bool pass = false;
        uint32_t j = 0;
        Cand c{};
        if (lane < cnt) {
            j = cand[lane];
            c = fetch(j);
            pass = near_group(c.p, 0.5f * c.h, gx0, gy0, gz0, gx1, gy1, gz1, qg);
        }
        const uint64_t sv = __ballot(pass);
        const int nc = __popcll(sv);
        if (pass) {
            const int at = prefix_count(sv);
            cand[at] = j;
            put(c, j, at);
            note_owner(own, j, grp.x, grp.y, at);
        }
        __syncthreads();
        pairs(nc, (split ? pair_mask_half(cP, nc, me, qi, lane >= kWave / 2) : pair_mask(cP, nc, me, qi)) &
                      ~self_bit(own, cand, nc, member, t));
    };
    // The density walk of this evaluation stored the group's candidates batch by batch: replay them in the
    // same order (so the sums are the walk's, bit for bit), or walk when the list did not fit.
    const uint32_t head = chead ? chead[gi] : kChunkOverflow;
    if (head == kChunkOverflow) {
        neighbour_walk(stk, cand, n, K, L, nS, expand, gx0, gy0, gz0, gx1, gy1, gz1, hg,
                       counters ? counters + (size_t)blockIdx.x * kCounterSlots + 11 : nullptr,
                       flush, sinst ? iroot[sinst[grp.x]] : 0u);
    } else if (head != kNoChunk) {  // (a group without candidates has no chunk)
        // The replay runs ahead of the pair terms: while chunk k is processed, the count, link, candidate
        // ids and masks of chunk k+2 are in flight and those of chunk k+1 have arrived, so a chunk waits for
        // one round trip (its particles' data). Loading that data a chunk ahead too measured no faster.
        // Every chunk owns 64 slots, so the slot reads past its count stay inside the pool (and go unused).
        uint32_t cnt = ccnt[head], nx = cnext[head];
        uint32_t cj = cpool[(size_t)head * kWave + lane];
        uint64_t m = cmask[(size_t)head * kWave + lane];
        uint32_t cnt1 = 0, nx1 = kNoChunk, cj1 = 0;
        uint64_t m1 = 0ull;
        if (nx != kNoChunk) {
            cnt1 = ccnt[nx];
            nx1 = cnext[nx];
            cj1 = cpool[(size_t)nx * kWave + lane];
            m1 = cmask[(size_t)nx * kWave + lane];
        }
        while (true) {
            if ((uint32_t)lane < cnt) put(fetch(cj), cj, lane);
            uint32_t cnt2 = 0, nx2 = kNoChunk, cj2 = 0;
            uint64_t m2 = 0ull;
            if (nx1 != kNoChunk) {
                cnt2 = ccnt[nx1];
                nx2 = cnext[nx1];
                cj2 = cpool[(size_t)nx1 * kWave + lane];
                m2 = cmask[(size_t)nx1 * kWave + lane];
            }
            __syncthreads();
            pairs((int)cnt, m);
            __syncthreads();
            if (nx == kNoChunk) break;
            cnt = cnt1;
            nx = nx1;
            cj = cj1;
            m = m1;
            cnt1 = cnt2;
            nx1 = nx2;
            cj1 = cj2;
            m1 = m2;
        }
    }
    if (split) {  // the upper half's sums join their member's lane
        if (F64) {
            ax64 += __shfl_down(ax64, kWave / 2);
            ay64 += __shfl_down(ay64, kWave / 2);
            az64 += __shfl_down(az64, kWave / 2);
            de64 += __shfl_down(de64, kWave / 2);
        } else {
            ax += __shfl_down(ax, kWave / 2);
            ay += __shfl_down(ay, kWave / 2);
            az += __shfl_down(az, kWave / 2);
            de += __shfl_down(de, kWave / 2);
        }
    }
    if (F64) {
        ax = (float)ax64;
        ay = (float)ay64;
        az = (float)az64;
        de = (float)de64;
    }
    if (valid && (!split || lane < kWave / 2)) {
        float4 gr = sgrav[t];
        // the un-resolved bulk and the external point masses (sph_step.wgsl:242-254), f32 as the shader
        float3 gb = bulk_gravity(make_float3(me.x, me.y, me.z), bnd.bulk_cr, bnd.bulk_vm);
        for (uint32_t k = 0; k < bnd.n_ext; ++k) {
            const float4 e = bnd.ext[k];
            const float dx = e.x - me.x, dy = e.y - me.y, dz = e.z - me.z;
            const float r2 = dx * dx + dy * dy + dz * dz + bnd.soft2;
            const float w = kG * e.w / (r2 * sqrtf(r2));
            gb.x += dx * w;
            gb.y += dy * w;
            gb.z += dz * w;
        }
        gr.x += gb.x;
        gr.y += gb.y;
        gr.z += gb.z;
        float4 out = me.w > 0.f ? make_float4(gr.x + ax, gr.y + ay, gr.z + az, de) : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t o = order[t];
        acc_orig[o] = out;
        if (F64 && acc64) {  // the f64 integrators take the sum unrounded
            const bool live = me.w > 0.f;
            if (REF) {  // the f64 gravity of the reference walk, plus the (f32) bulk and external terms
                const double4 g64 = sgrav64[t];
                acc64[3 * o] = live ? g64.x + (double)gb.x + ax64 : 0.0;
                acc64[3 * o + 1] = live ? g64.y + (double)gb.y + ay64 : 0.0;
                acc64[3 * o + 2] = live ? g64.z + (double)gb.z + az64 : 0.0;
            } else {
                // The f32 gravity carries G rounded to f32, 3.6e-8 too strong: a systematic force scale whose
                // work over a run shows as energy drift against the f64 energy. The f64 sum undoes it.
                constexpr double kGfix = kG64 / (double)kG;
                acc64[3 * o] = live ? (double)gr.x * kGfix + ax64 : 0.0;
                acc64[3 * o + 1] = live ? (double)gr.y * kGfix + ay64 : 0.0;
                acc64[3 * o + 2] = live ? (double)gr.z * kGfix + az64 : 0.0;
            }
        }
    }
    tested_all += tested * grp.y;
    }
    if (lane == 0 && counters) counters[(size_t)blockIdx.x * kCounterSlots + 4] += tested_all;
}

// ---- KDK (sph_step.wgsl:304-333; hydrostatic.rs:451-465), retired particles are not integrated ----------
// The half-kick of u clamps at zero (sph_step.wgsl:310, hydrostatic.rs:457): the energy that adds, m (-u_new)
// wherever u_new < 0, is counted in f64. Clamps are common in expanding material, so each 256-thread block
// adds its sum into its own slot (the two kicks of a step run one after the other, so no slot is written
// concurrently) instead of all blocks meeting on one address; the host sums the slots when asked. Every
// thread of the block reaches the reduction.
constexpr int kKickThreads = 256;
__device__ __forceinline__ float clamp_u(float u_new, float m, double& injected) {
    if (u_new < 0.0f) injected += -(double)u_new * (double)m;
    return fmaxf(u_new, 0.0f);
}
__device__ __forceinline__ void add_injected(double* inject, double injected) {
    __shared__ double part[kKickThreads / kWave];
    if (!inject || !__syncthreads_or(injected != 0.0)) return;
    const double s = wave_sum_f64(injected);
    if (lane_id() == 0) part[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < kKickThreads / kWave; ++w) b += part[w];
        inject[blockIdx.x] += b;
    }
}

__global__ __launch_bounds__(256) void k_kick_drift(uint32_t n, float4* posh, float4* velu,
                                                    const float* __restrict__ mass, const float4* __restrict__ acc,
                                                    const float* __restrict__ dtp, double* pos64, double* vel64,
                                                    const double* __restrict__ acc64,
                                                    const uint32_t* __restrict__ inst, const float* __restrict__ idt,
                                                    float4 bulk_cr, float4 bulk_vm, double* inject) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double injected = 0.0;
    if (i < n && mass[i] > 0.f) {
    const float dt = inst ? idt[inst[i]] : *dtp;
    const float4 a = acc[i];
    float4 v = velu[i];
    float4 p = posh[i];
    v.w = clamp_u(v.w + a.w * (0.5f * dt), mass[i], injected);
    if (pos64) {
        double hd = 0.5 * (double)dt;
        double vx = vel64[3 * i] + acc64[3 * i] * hd, vy = vel64[3 * i + 1] + acc64[3 * i + 1] * hd,
               vz = vel64[3 * i + 2] + acc64[3 * i + 2] * hd;
        double x = pos64[3 * i] + vx * (double)dt, y = pos64[3 * i + 1] + vy * (double)dt,
               z = pos64[3 * i + 2] + vz * (double)dt;
        bulk_floor(x, y, z, vx, vy, vz, bulk_cr, bulk_vm);  // sph_step.wgsl:319
        vel64[3 * i] = vx;
        vel64[3 * i + 1] = vy;
        vel64[3 * i + 2] = vz;
        pos64[3 * i] = x;
        pos64[3 * i + 1] = y;
        pos64[3 * i + 2] = z;
        v.x = (float)vx;
        v.y = (float)vy;
        v.z = (float)vz;
        p.x = (float)x;
        p.y = (float)y;
        p.z = (float)z;
    } else {
        v.x += a.x * (0.5f * dt);
        v.y += a.y * (0.5f * dt);
        v.z += a.z * (0.5f * dt);
        p.x += v.x * dt;
        p.y += v.y * dt;
        p.z += v.z * dt;
        bulk_floor(p.x, p.y, p.z, v.x, v.y, v.z, bulk_cr, bulk_vm);  // sph_step.wgsl:319
    }
    velu[i] = v;
    posh[i] = p;
    }
    add_injected(inject, injected);
}

// Damped relaxation (sph_step.wgsl:340-359): v = (v + (a + w^2 (x, y, 0)) dt) * damp, x += v dt, floor.
// Internal energy is held. One relaxation step is one force evaluation and this kernel.
__global__ __launch_bounds__(256) void k_relax(uint32_t n, float4* posh, float4* velu, const float* __restrict__ mass,
                                               const float4* __restrict__ acc, const float* __restrict__ dtp,
                                               float damp, float omega, double* pos64, double* vel64,
                                               const double* __restrict__ acc64, float4 bulk_cr, float4 bulk_vm) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(mass[i] > 0.f)) return;
    const float dt = *dtp;
    float4 v = velu[i];
    float4 p = posh[i];
    const float w2 = omega * omega;
    if (pos64) {
        double x = pos64[3 * i], y = pos64[3 * i + 1], z = pos64[3 * i + 2];
        double vx = (vel64[3 * i] + (acc64[3 * i] + (double)w2 * x) * dt) * damp;
        double vy = (vel64[3 * i + 1] + (acc64[3 * i + 1] + (double)w2 * y) * dt) * damp;
        double vz = (vel64[3 * i + 2] + acc64[3 * i + 2] * dt) * damp;
        x += vx * dt;
        y += vy * dt;
        z += vz * dt;
        bulk_floor(x, y, z, vx, vy, vz, bulk_cr, bulk_vm);
        pos64[3 * i] = x;
        pos64[3 * i + 1] = y;
        pos64[3 * i + 2] = z;
        vel64[3 * i] = vx;
        vel64[3 * i + 1] = vy;
        vel64[3 * i + 2] = vz;
        p.x = (float)x;
        p.y = (float)y;
        p.z = (float)z;
        v.x = (float)vx;
        v.y = (float)vy;
        v.z = (float)vz;
    } else {
        const float4 a = acc[i];
        v.x = (v.x + (a.x + w2 * p.x) * dt) * damp;
        v.y = (v.y + (a.y + w2 * p.y) * dt) * damp;
        v.z = (v.z + a.z * dt) * damp;
        p.x += v.x * dt;
        p.y += v.y * dt;
        p.z += v.z * dt;
        bulk_floor(p.x, p.y, p.z, v.x, v.y, v.z, bulk_cr, bulk_vm);
    }
    velu[i] = v;
    posh[i] = p;
}

__global__ __launch_bounds__(256) void k_kick(uint32_t n, float4* velu, const float* __restrict__ mass,
                                              const float4* __restrict__ acc, const float* __restrict__ dtp,
                                              double* vel64, const double* __restrict__ acc64,
                                              const uint32_t* __restrict__ inst, const float* __restrict__ idt,
                                              double* inject) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double injected = 0.0;
    if (i < n && mass[i] > 0.f) {
    const float dt = inst ? idt[inst[i]] : *dtp;
    const float4 a = acc[i];
    float4 v = velu[i];
    v.w = clamp_u(v.w + a.w * (0.5f * dt), mass[i], injected);
    if (vel64) {
        double hd = 0.5 * (double)dt;
        double vx = vel64[3 * i] + acc64[3 * i] * hd, vy = vel64[3 * i + 1] + acc64[3 * i + 1] * hd,
               vz = vel64[3 * i + 2] + acc64[3 * i + 2] * hd;
        vel64[3 * i] = vx;
        vel64[3 * i + 1] = vy;
        vel64[3 * i + 2] = vz;
        v.x = (float)vx;
        v.y = (float)vy;
        v.z = (float)vz;
    } else {
        v.x += a.x * (0.5f * dt);
        v.y += a.y * (0.5f * dt);
        v.z += a.z * (0.5f * dt);
    }
    velu[i] = v;
    }
    add_injected(inject, injected);
}

// dt = cfl * min_i h_i / max(c_i + |v_i|, 1)  (impact-run main.rs:345-356 done on the device).
__global__ void k_set_dt_cfl(const uint32_t* scal, float cfl, float* dt) {
    float m = __uint_as_float(scal[6]);
    *dt = (m < INFINITY) ? cfl * m : 0.0f;
}
__global__ void k_set_dt_value(float v, float* dt) { *dt = v; }
__global__ void k_advance_time(const float* dt, double* time) { *time += (double)*dt; }

__global__ __launch_bounds__(256) void k_unpack(uint32_t n, const ParticleRec* __restrict__ rec, float4* posh,
                                                float4* velu, float* mass, uint32_t* mat, uint32_t* prov, float* rho) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ParticleRec r = rec[i];
    posh[i] = make_float4(r.px, r.py, r.pz, r.h);
    velu[i] = make_float4(r.vx, r.vy, r.vz, r.u);
    mass[i] = r.mass;
    mat[i] = r.mat;
    prov[i] = r.prov;
    rho[i] = r.rho;
}
__global__ __launch_bounds__(256) void k_pack(uint32_t n, ParticleRec* __restrict__ rec, const float4* posh,
                                              const float4* velu, const float* mass, const uint32_t* mat,
                                              const uint32_t* prov, const float* rho) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = posh[i], v = velu[i];
    ParticleRec r;
    r.px = p.x; r.py = p.y; r.pz = p.z; r.h = p.w;
    r.vx = v.x; r.vy = v.y; r.vz = v.z; r.u = v.w;
    r.mass = mass[i];
    r.mat = mat[i];
    r.rho = rho[i];
    r.prov = prov[i];
    rec[i] = r;
}
__global__ __launch_bounds__(256) void k_sync64(uint32_t n, const float4* posh, const float4* velu, double* pos64,
                                                double* vel64) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = posh[i], v = velu[i];
    pos64[3 * i] = p.x; pos64[3 * i + 1] = p.y; pos64[3 * i + 2] = p.z;
    vel64[3 * i] = v.x; vel64[3 * i + 1] = v.y; vel64[3 * i + 2] = v.z;
}

// ---- f64 diagnostics: per-block partials, summed on the host in a fixed order ------------------------------
constexpr int kDiagBlocks = 1024;
constexpr int kDiagFields = 16;

__device__ __forceinline__ void load_pos64(uint32_t i, const float4* posh, const double* pos64, double* x) {
    if (pos64) {
        for (int k = 0; k < 3; ++k) x[k] = pos64[3 * i + k];
    } else {
        float4 p = posh[i];
        x[0] = p.x; x[1] = p.y; x[2] = p.z;
    }
}

__device__ __forceinline__ void load_state64(uint32_t i, const float4* posh, const float4* velu, const double* pos64,
                                             const double* vel64, double* x, double* v) {
    if (pos64) {
        for (int k = 0; k < 3; ++k) {
            x[k] = pos64[3 * i + k];
            v[k] = vel64[3 * i + k];
        }
    } else {
        float4 p = posh[i], q = velu[i];
        x[0] = p.x; x[1] = p.y; x[2] = p.z;
        v[0] = q.x; v[1] = q.y; v[2] = q.z;
    }
}

// Fields: 0 KE, 1 IE, 2 mass, 3-5 sum m x, 6-8 sum m v, 9-11 sum m (x cross v), 12 PE partial.
__global__ __launch_bounds__(256) void k_diag_sums(uint32_t n, const float4* posh, const float4* velu,
                                                   const double* pos64, const double* vel64, const float* mass,
                                                   double* red) {
    __shared__ double part[4][12];
    double acc[12];
    for (int k = 0; k < 12; ++k) acc[k] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double x[3], v[3];
        load_state64(i, posh, velu, pos64, vel64, x, v);
        double m = mass[i];
        double u = velu[i].w;
        acc[0] += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        acc[1] += m * u;
        acc[2] += m;
        for (int k = 0; k < 3; ++k) {
            acc[3 + k] += m * x[k];
            acc[6 + k] += m * v[k];
        }
        acc[9] += m * (x[1] * v[2] - x[2] * v[1]);
        acc[10] += m * (x[2] * v[0] - x[0] * v[2]);
        acc[11] += m * (x[0] * v[1] - x[1] * v[0]);
    }
    int w = threadIdx.x / kWave;
    for (int k = 0; k < 12; ++k) {
        double s = wave_sum_f64(acc[k]);
        if (lane_id() == 0) part[w][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        double s = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
        red[blockIdx.x * kDiagFields + threadIdx.x] = s;
    }
}

// Softened potential energy, exact O(N^2) in f64: PE = -G/2 sum_i m_i sum_{j != i} m_j / sqrt(r^2 + s^2)
// (gpu_sph.rs:940-961 sums i<j on the CPU). 1/sqrt by an f32 seed plus two Newton steps in f64.
__global__ __launch_bounds__(256) void k_diag_pe(uint32_t n, const float4* posh, const double* pos64,
                                                 const float* mass, double soft2, double* red) {
    __shared__ double tx[256], ty[256], tz[256], tm[256];
    double pe_block = 0.0;
    for (uint32_t ib = blockIdx.x * 256; ib < n; ib += gridDim.x * 256) {
        uint32_t i = ib + threadIdx.x;
        double xi[3] = {0, 0, 0};
        double mi = 0.0;
        if (i < n) {
            load_pos64(i, posh, pos64, xi);
            mi = mass[i];
        }
        double phi = 0.0;
        for (uint32_t jb = 0; jb < n; jb += 256) {
            uint32_t j = jb + threadIdx.x;
            __syncthreads();
            if (j < n) {
                double xj[3];
                load_pos64(j, posh, pos64, xj);
                tx[threadIdx.x] = xj[0];
                ty[threadIdx.x] = xj[1];
                tz[threadIdx.x] = xj[2];
                tm[threadIdx.x] = mass[j];
            } else {
                tx[threadIdx.x] = ty[threadIdx.x] = tz[threadIdx.x] = 0.0;
                tm[threadIdx.x] = 0.0;
            }
            __syncthreads();
            uint32_t cnt = min(256u, n - jb);
            for (uint32_t jj = 0; jj < cnt; ++jj) {
                double dx = tx[jj] - xi[0], dy = ty[jj] - xi[1], dz = tz[jj] - xi[2];
                double r2 = dx * dx + dy * dy + dz * dz + soft2;
                double y = (double)__builtin_amdgcn_rsqf((float)r2);
                y = y * (1.5 - 0.5 * r2 * y * y);
                y = y * (1.5 - 0.5 * r2 * y * y);
                phi += (jb + jj == i) ? 0.0 : tm[jj] * y;
            }
        }
        pe_block += mi * phi;
    }
    __shared__ double part[4];
    double s = wave_sum_f64(pe_block);
    if (lane_id() == 0) part[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) red[blockIdx.x * kDiagFields + 12] = -0.5 * kG64 * (part[0] + part[1] + part[2] + part[3]);
}

static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

static CandCache cand_cache(const Dev& d, const StepParams& p) {
    return CandCache{d.chead, d.cnext, d.ccnt, d.cpool, d.cmask, std::min(d.cblocks, p.cache_blocks)};
}

static Boundary boundary(const Dev& d, const StepParams& p) {
    return Boundary{p.bulk_cr, p.bulk_vm, d.ext, p.n_ext, p.soft2};
}

void launch_density(const Dev& d, const StepParams& p, hipStream_t s) {
    uint32_t groups = std::min<uint32_t>(kWalkWaves, d.n);
    if (p.ref64)
        hipLaunchKernelGGL((k_density<true, true>), dim3(groups), dim3(kWave), 0, s, d.spos, d.sh, d.n, d.K, d.L,
                           d.nS, d.vals, d.srho, d.rho, d.srho64, d.scal, d.counters, d.groups, d.nrange,
                           (uint32_t)p.sph_expand, d.sinst(), d.iroot, cand_cache(d, p), (const float4*)d.slo);
    else if (p.mixed)
        hipLaunchKernelGGL(k_density<true>, dim3(groups), dim3(kWave), 0, s, d.spos, d.sh, d.n, d.K, d.L, d.nS,
                           d.vals, d.srho, d.rho, d.srho64, d.scal, d.counters, d.groups, d.nrange,
                           (uint32_t)p.sph_expand, d.sinst(), d.iroot, cand_cache(d, p), (const float4*)nullptr);
    else
        hipLaunchKernelGGL(k_density<false>, dim3(groups), dim3(kWave), 0, s, d.spos, d.sh, d.n, d.K, d.L, d.nS,
                           d.vals, d.srho, d.rho, (double*)nullptr, d.scal, d.counters, d.groups,
                           d.nrange, (uint32_t)p.sph_expand, d.sinst(), d.iroot, cand_cache(d, p),
                           (const float4*)nullptr);
    launch_eos(d, s);
}
void launch_eos(const Dev& d, hipStream_t s) {
    const uint32_t blocks = std::min<uint32_t>(kEosBlocks, cdiv(d.n, kEosThreads));
    hipLaunchKernelGGL(k_eos, dim3(blocks), dim3(kEosThreads), 0, s, d.sh, d.svel, d.smat, d.n, d.eos, d.srho, d.spr,
                       d.scs, (const double*)d.srho64, d.spr64, d.scs64, d.scal, d.sinst(), d.isig);
}
void launch_force(const Dev& d, const StepParams& p, hipStream_t s) {
    uint32_t groups = std::min<uint32_t>(kWalkWaves, d.n);
    if (p.ref64)
        hipLaunchKernelGGL((k_force<true, true>), dim3(groups), dim3(kWave), 0, s, d.spos, d.sh, d.svel, d.srho,
                           d.spr, d.scs, d.n, d.K, d.L, d.nS, d.vals, d.sgrav, d.acc, d.srho64, d.spr64,
                           d.scs64, p.av_alpha, p.av_beta, d.counters, d.groups, d.scal, d.nrange,
                           (uint32_t)p.sph_expand, d.acc64, d.sinst(), d.iroot, boundary(d, p), d.chead, d.cnext,
                           d.ccnt, d.cpool, d.cmask, (const float4*)d.slo, (const float4*)d.svlo,
                           (const double4*)d.sgrav64);
    else if (p.mixed)
        hipLaunchKernelGGL(k_force<true>, dim3(groups), dim3(kWave), 0, s, d.spos, d.sh, d.svel, d.srho, d.spr, d.scs,
                           d.n, d.K, d.L, d.nS, d.vals, d.sgrav, d.acc, d.srho64, d.spr64,
                           d.scs64, p.av_alpha, p.av_beta, d.counters, d.groups, d.scal, d.nrange,
                           (uint32_t)p.sph_expand, d.acc64, d.sinst(), d.iroot, boundary(d, p), d.chead, d.cnext,
                           d.ccnt, d.cpool, d.cmask, (const float4*)nullptr, (const float4*)nullptr,
                           (const double4*)nullptr);
    else
        hipLaunchKernelGGL(k_force<false>, dim3(groups), dim3(kWave), 0, s, d.spos, d.sh, d.svel, d.srho, d.spr,
                           d.scs, d.n, d.K, d.L, d.nS, d.vals, d.sgrav, d.acc,
                           (const double*)nullptr, (const double*)nullptr, (const double*)nullptr, p.av_alpha,
                           p.av_beta, d.counters, d.groups, d.scal, d.nrange, (uint32_t)p.sph_expand,
                           (double*)nullptr, d.sinst(), d.iroot, boundary(d, p), d.chead, d.cnext, d.ccnt, d.cpool,
                           d.cmask, (const float4*)nullptr, (const float4*)nullptr, (const double4*)nullptr);
}
void launch_kick_drift(const Dev& d, const StepParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_kick_drift, dim3(cdiv(d.n, 256)), dim3(256), 0, s, d.n, d.posh, d.velu, d.mass, d.acc, d.dt,
                       d.pos64, d.vel64, d.acc64, d.ninst > 1 ? d.inst : nullptr, d.idt, p.bulk_cr, p.bulk_vm,
                       d.inject);
}
void launch_relax(const Dev& d, const StepParams& p, float damp, hipStream_t s) {
    hipLaunchKernelGGL(k_relax, dim3(cdiv(d.n, 256)), dim3(256), 0, s, d.n, d.posh, d.velu, d.mass, d.acc, d.dt,
                       damp, p.omega, d.pos64, d.vel64, d.acc64, p.bulk_cr, p.bulk_vm);
}
void launch_kick(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_kick, dim3(cdiv(d.n, 256)), dim3(256), 0, s, d.n, d.velu, d.mass, d.acc, d.dt, d.vel64,
                       d.acc64, d.ninst > 1 ? d.inst : nullptr, d.idt, d.inject);
}
void launch_set_dt_cfl(const Dev& d, float cfl, hipStream_t s) {
    hipLaunchKernelGGL(k_set_dt_cfl, dim3(1), dim3(1), 0, s, d.scal, cfl, d.dt);
}
void launch_set_dt_value(const Dev& d, float dt, hipStream_t s) {
    hipLaunchKernelGGL(k_set_dt_value, dim3(1), dim3(1), 0, s, dt, d.dt);
}
void launch_advance_time(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, s, d.dt, d.time);
}
void launch_unpack(const Dev& d, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_unpack, dim3(cdiv(n, 256)), dim3(256), 0, s, n, (const ParticleRec*)d.aos, d.posh, d.velu,
                       d.mass, d.mat, d.prov, d.rho);
}
void launch_pack(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_pack, dim3(cdiv(d.n, 256)), dim3(256), 0, s, d.n, (ParticleRec*)d.aos, d.posh, d.velu, d.mass,
                       d.mat, d.prov, d.rho);
}
void launch_sync64_from32(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_sync64, dim3(cdiv(d.n, 256)), dim3(256), 0, s, d.n, d.posh, d.velu, d.pos64, d.vel64);
}
void launch_diag(const Dev& d, float soft, int with_pe, hipStream_t s) {
    hipLaunchKernelGGL(k_diag_sums, dim3(kDiagBlocks), dim3(256), 0, s, d.n, d.posh, d.velu, d.pos64, d.vel64, d.mass,
                       d.red);
    if (with_pe) {
        uint32_t blocks = std::min<uint32_t>(cdiv(d.n, 256), kDiagBlocks);
        hipLaunchKernelGGL(k_diag_pe, dim3(blocks), dim3(256), 0, s, d.n, d.posh, d.pos64, d.mass,
                           (double)soft * (double)soft, d.red);
    }
}

}  // namespace isph
