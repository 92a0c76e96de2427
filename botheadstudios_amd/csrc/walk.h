// Wave-cooperative neighbour search over the LBVH, shared by the SPH walks and their adjoints.
#pragma once

#include "state.h"

namespace isph {

#ifndef ISPH_SPH_STACK
#define ISPH_SPH_STACK 512
#endif
constexpr int kSphStack = ISPH_SPH_STACK;  // LDS traversal stack (entries); the pop width narrows near it
constexpr int kSphStackMargin = 128;       // > maximum Karras depth
constexpr int kSphBuf = 128;               // candidate list (two flushes of 64)

// Wave-cooperative neighbour traversal from node `root` (the whole tree, or an ensemble instance's
// subtree). `flush(cnt)` is called with cand[0..cnt) holding candidate sorted indices (cnt <= 64); it must
// evaluate them for the calling lane.
template <class Flush>
__device__ __forceinline__ void neighbour_walk(uint32_t* stk, uint32_t* cand, uint32_t n, uint32_t K, uint32_t L,
                                               const SphNode* __restrict__ nS, uint32_t expand,
                                               float gx0, float gy0, float gz0, float gx1, float gy1, float gz1,
                                               float hgmax, unsigned long long* stats, Flush&& flush,
                                               uint32_t root = 0) {
    const int lane = threadIdx.x;
    const uint32_t leaf0 = L - 1;
    int sp = 1, nq = 0;
    unsigned long long rounds = 0, popped = 0;
    int max_sp = 1;
    if (lane == 0) stk[0] = root;
    __syncthreads();
    auto shift = [&](int cnt) {
        __syncthreads();
        if (lane + kWave < cnt) cand[lane] = cand[lane + kWave];
        __syncthreads();
    };
    while (sp > 0) {
        int k = min(kWave, sp);
        int room = kSphStack - kSphStackMargin - sp;
        if (room < k) k = max(1, room);
        sp -= k;
        rounds += 1;
        popped += k;
        const bool has = lane < k;
        const uint32_t node = has ? stk[sp + lane] : 0u;
        __syncthreads();
        bool keep = false;
        uint4 ch = make_uint4(0u, 0u, 0u, 0u);  // children and leaf range travel with the box: one memory
        if (has) {                               // round trip per round
            const float4 B = nS[node].lo;
            const float4 C = nS[node].hi;
            if (node < leaf0) ch = nS[node].lr;
            float dx = fmaxf(fmaxf(B.x - gx1, gx0 - C.x), 0.f);
            float dy = fmaxf(fmaxf(B.y - gy1, gy0 - C.y), 0.f);
            float dz = fmaxf(fmaxf(B.z - gz1, gz0 - C.z), 0.f);
            float rr = 0.5f * (hgmax + B.w);
            keep = dx * dx + dy * dy + dz * dz < rr * rr;
        }
        // A kept subtree of <= `expand` particles is taken whole (its leaf range is contiguous in sorted
        // order) instead of visiting its leaves one round at a time; the exact radius test follows anyway.
        uint32_t first = 0, cnt = 0;
        if (keep) {
            if (node >= leaf0) {
                first = (node - leaf0) * K;
                cnt = min(first + K, n) - first;
            } else {
                const uint32_t pf = ch.z * K, pe = min((ch.w + 1) * K, n);
                if (pe - pf <= expand) {
                    first = pf;
                    cnt = pe - pf;
                }
            }
        }
        const bool leaf = keep && cnt > 0;
        const bool open = keep && cnt == 0;
        const uint64_t mo = __ballot(open);
        if (open) {
            int pos = sp + 2 * prefix_count(mo);
            stk[pos] = ch.x;
            stk[pos + 1] = ch.y;
        }
        sp += 2 * __popcll(mo);
        max_sp = max(max_sp, sp);
        const uint64_t ml = __ballot(leaf);
        if (ml) {
            const uint32_t width = max(K, expand);
            for (uint32_t kk = 0; kk < width; ++kk) {
                const bool present = leaf && kk < cnt;
                const uint64_t mp = __ballot(present);
                if (!mp) break;
                if (present) cand[nq + prefix_count(mp)] = first + kk;
                nq += __popcll(mp);
                if (nq >= kWave) {
                    __syncthreads();
                    flush(kWave);
                    shift(nq);
                    nq -= kWave;
                }
            }
        }
        __syncthreads();
    }
    if (nq > 0) {
        __syncthreads();
        flush(nq);
        __syncthreads();
    }
    if (lane == 0 && stats) {
        stats[0] += rounds;  // the calling wave's own slots
        stats[1] += popped;
        stats[2] = max(stats[2], (unsigned long long)max_sp);
    }
}

// The pair test of one lane against the candidates in LDS (x, y, z, h/2 per entry): bit jj is set when
// r < (h_i + h_j) / 2 (sph_step.wgsl:201,210 / :262-272), one test for the density and the force passes of
// both precision forms (the force skips coincident pairs, r <= 1e-9, inside its pair loop). The lane's own
// particle is removed afterwards with `self_bit`. Halving is exact, so h_i/2 + h_j/2 is bit for bit the
// reference's (h_i + h_j) * 0.5. Each half of the mask is built from its highest slot down: the bit is the
// sign of r^2 - h_ij^2 (negative, or -0 after a flush to zero, exactly when r^2 < h_ij^2), shifted in with
// one funnel shift per candidate.
__device__ __forceinline__ uint32_t pair_miss_sign(float4 pj, float4 me, float qi) {
    const float dx = me.x - pj.x, dy = me.y - pj.y, dz = me.z - pj.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float hij = qi + pj.w;
    float d;
    {
#pragma clang fp contract(off)
        const float hh = hij * hij;  // h_ij^2 rounded on its own, never fused into r^2 - h_ij^2
        d = r2 - hh;
    }
    return __float_as_uint(d);
}
__device__ __forceinline__ uint64_t pair_mask(const float4* cP, int cnt, float4 me, float qi) {
    uint32_t hi = 0u, lo = 0u;
#pragma unroll 4
    for (int jj = min(cnt, kWave) - 1; jj >= 32; --jj)
        hi = __builtin_amdgcn_alignbit(hi, pair_miss_sign(cP[jj], me, qi), 31u);
#pragma unroll 4
    for (int jj = min(cnt, 32) - 1; jj >= 0; --jj)
        lo = __builtin_amdgcn_alignbit(lo, pair_miss_sign(cP[jj], me, qi), 31u);
    return ((uint64_t)hi << 32) | lo;
}

// The candidate prefilter of the SPH walks: a candidate (x, y, z, h_j / 2) can pair with a member of the
// group only if its distance to the group box is below (max h_i + h_j) / 2. The walk keeps whole subtrees of
// up to `expand` particles, so some candidates fail it. The margin (1e-4 of the squared radius) is far above
// the few ulps by which the f32 radius test and this bound can differ, so no pair is lost. The density walk
// and the force pass's own walk drop the same candidates and pack the rest in walk order, so both see the
// same batches.
__device__ __forceinline__ bool near_group(float4 p, float qj, float gx0, float gy0, float gz0, float gx1,
                                           float gy1, float gz1, float qg) {
    const float dx = fmaxf(fmaxf(gx0 - p.x, p.x - gx1), 0.f);
    const float dy = fmaxf(fmaxf(gy0 - p.y, p.y - gy1), 0.f);
    const float dz = fmaxf(fmaxf(gz0 - p.z, p.z - gz1), 0.f);
    const float rr = qg + qj;
    return dx * dx + dy * dy + dz * dz <= rr * rr * (1.0f + 1.0e-4f);
}

// While loading candidate `cand[lane]`, record where a particle of the lane's own group sits in this batch
// (own[d] = slot for group member d). A lane then finds its own particle with one LDS read and a check,
// instead of comparing indices at every candidate. Stale entries from earlier batches fail the check.
__device__ __forceinline__ void note_owner(uint32_t* own, uint32_t j, uint32_t g0, uint32_t gn, int lane) {
    const uint32_t d = j - g0;
    if (d < gn) own[d] = (uint32_t)lane;
}
__device__ __forceinline__ uint64_t self_bit(const uint32_t* own, const uint32_t* cand, int cnt, int member,
                                             uint32_t t) {
    const uint32_t ms = own[member];
    return (ms < (uint32_t)cnt && cand[ms] == t) ? (1ull << ms) : 0ull;
}
// The split form for a group of at most 32 particles held by both wave halves (group member = lane mod
// 32): the lower half tests slots 0..31 and the upper half slots 32..63 of the batch, each bit landing in
// its own half of the mask, so a batch costs half the steps; the two lanes of a member later add their
// sums. The test and its bits are pair_mask's.
__device__ __forceinline__ uint64_t pair_mask_half(const float4* cP, int cnt, float4 me, float qi, bool upper) {
    const int base = upper ? 32 : 0, h = min(cnt, 32);
    uint32_t bits = 0u;
#pragma unroll 4
    for (int jj = h - 1; jj >= 0; --jj) {
        const bool in = jj + base < cnt;
        const uint32_t sgn = pair_miss_sign(cP[in ? jj + base : jj], me, qi);
        bits = __builtin_amdgcn_alignbit(bits, in ? sgn : 0u, 31u);
    }
    return upper ? ((uint64_t)bits << 32) : (uint64_t)bits;
}

}  // namespace isph
