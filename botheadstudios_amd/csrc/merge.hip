// Pairwise merging under a particle budget (sph_step.wgsl:374-541): three passes with disjoint write sets.
//
// pick:   every live particle chooses the neighbour that will absorb it: same material, strictly "greater"
//         (heavier, or equal mass and lower upload index), closer than half the pair's mean smoothing length
//         and moving subsonically relative to it (|v_i - v_j| < c_i); the heaviest such neighbour wins, the
//         lowest upload index among equally heavy ones (so the choice does not depend on the walk order).
// apply:  every root (a particle that picked nobody) absorbs the neighbours that picked it: mass, momentum
//         and centre of mass add up, the destroyed relative kinetic energy becomes internal energy, and h
//         grows as (m_new / m)^(1/3).
// retire: a particle whose absorber was a root becomes inert (mass 0, velocity 0).
// Chains (i -> j -> k) resolve over successive passes. Neighbours come from the LBVH walk (a superset of the
// pairs within the pair's mean h), the sound speed from the EOS pass of the density evaluation.
#include "walk.h"

namespace isph {

__global__ __launch_bounds__(64) void k_merge_pick(const float4* __restrict__ spos, const float* __restrict__ sh,
                                                   const float4* __restrict__ svel, const uint32_t* __restrict__ smat,
                                                   const float* __restrict__ scs, uint32_t n, uint32_t K, uint32_t L,
                                                   const SphNode* __restrict__ nS, const uint2* __restrict__ nrange,
                                                   uint32_t expand, const uint32_t* __restrict__ order,
                                                   const uint2* __restrict__ groups, const uint32_t* __restrict__ scal,
                                                   const uint32_t* __restrict__ sinst,
                                                   const uint32_t* __restrict__ iroot, uint32_t* __restrict__ target) {
    __shared__ uint32_t stk[kSphStack];
    __shared__ uint32_t cand[kSphBuf];
    const int lane = threadIdx.x;
    const uint32_t ngroups = scal[kGroupsSph];
    for (uint32_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const uint2 grp = groups[gi];
        const uint32_t t = grp.x + lane;
        const bool valid = (uint32_t)lane < grp.y;
        const uint32_t tc = valid ? t : grp.x;
        __syncthreads();
        const float4 me = spos[tc];
        const float hi = sh[tc];
        const float4 vi = svel[tc];
        const uint32_t mi = smat[tc];
        const float ci = scs[tc];
        const uint32_t oi = order[tc];
        const bool live = valid && me.w > 0.f;
        const float gx0 = wave_min(me.x), gy0 = wave_min(me.y), gz0 = wave_min(me.z);
        const float gx1 = wave_max(me.x), gy1 = wave_max(me.y), gz1 = wave_max(me.z);
        const float hg = wave_max(hi);
        uint32_t best = kNoMerge;
        float best_mass = me.w;
        neighbour_walk(stk, cand, n, K, L, nS, expand, gx0, gy0, gz0, gx1, gy1, gz1, hg, nullptr,
                       [&](int cnt) {
            for (int jj = 0; jj < cnt; ++jj) {
                const uint32_t j = cand[jj];
                if (!live || j == t) continue;
                const float4 pj = spos[j];
                if (!(pj.w > 0.f) || smat[j] != mi) continue;  // same material only (:441-447)
                const uint32_t oj = order[j];
                const bool greater = pj.w > me.w || (pj.w == me.w && oj < oi);
                if (!greater) continue;
                const float hij = 0.5f * (hi + sh[j]);
                const float dx = me.x - pj.x, dy = me.y - pj.y, dz = me.z - pj.z;
                if (!(sqrtf(dx * dx + dy * dy + dz * dz) < 0.5f * hij)) continue;
                const float4 vj = svel[j];
                const float ux = vi.x - vj.x, uy = vi.y - vj.y, uz = vi.z - vj.z;
                if (!(sqrtf(ux * ux + uy * uy + uz * uz) < ci)) continue;
                // the heaviest wins; among equally heavy absorbers the lowest upload index (the reference
                // keeps the first in its hash-bucket order, which is not deterministic)
                if (best == kNoMerge || pj.w > best_mass || (pj.w == best_mass && oj < best)) {
                    best = oj;
                    best_mass = pj.w;
                }
            }
        }, sinst ? iroot[sinst[grp.x]] : 0u);
        if (valid) target[oi] = best;
    }
}

__global__ __launch_bounds__(64) void k_merge_apply(const float4* __restrict__ spos, const float* __restrict__ sh,
                                                    const float4* __restrict__ svel, uint32_t n, uint32_t K,
                                                    uint32_t L, const SphNode* __restrict__ nS,
                                                    const uint2* __restrict__ nrange, uint32_t expand,
                                                    const uint32_t* __restrict__ order,
                                                    const uint2* __restrict__ groups,
                                                    const uint32_t* __restrict__ scal,
                                                    const uint32_t* __restrict__ sinst,
                                                    const uint32_t* __restrict__ iroot,
                                                    const uint32_t* __restrict__ target, float4* posh, float4* velu,
                                                    float* mass, double* pos64, double* vel64) {
    __shared__ uint32_t stk[kSphStack];
    __shared__ uint32_t cand[kSphBuf];
    const int lane = threadIdx.x;
    const uint32_t ngroups = scal[kGroupsSph];
    for (uint32_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const uint2 grp = groups[gi];
        const uint32_t t = grp.x + lane;
        const bool valid = (uint32_t)lane < grp.y;
        const uint32_t tc = valid ? t : grp.x;
        __syncthreads();
        const float4 me = spos[tc];
        const float hj = sh[tc];
        const uint32_t oj = order[tc];
        const bool root = valid && me.w > 0.f && target[oj] == kNoMerge;
        const float gx0 = wave_min(me.x), gy0 = wave_min(me.y), gz0 = wave_min(me.z);
        const float gx1 = wave_max(me.x), gy1 = wave_max(me.y), gz1 = wave_max(me.z);
        const float hg = wave_max(hj);
        // the root itself, then every neighbour that picked it (f64 sums; f64 state in the mixed form)
        auto px = [&](uint32_t o, uint32_t s, int k) -> double {
            return pos64 ? pos64[3 * o + k] : (double)(k == 0 ? spos[s].x : (k == 1 ? spos[s].y : spos[s].z));
        };
        auto vx = [&](uint32_t o, uint32_t s, int k) -> double {
            return vel64 ? vel64[3 * o + k] : (double)(k == 0 ? svel[s].x : (k == 1 ? svel[s].y : svel[s].z));
        };
        double m = me.w, mom[3], com[3], mu = (double)me.w * svel[tc].w, ke = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double v = vx(oj, tc, k);
            mom[k] = m * v;
            com[k] = m * px(oj, tc, k);
            ke += 0.5 * m * v * v;
        }
        uint32_t absorbed = 0;
        neighbour_walk(stk, cand, n, K, L, nS, expand, gx0, gy0, gz0, gx1, gy1, gz1, hg, nullptr,
                       [&](int cnt) {
            for (int jj = 0; jj < cnt; ++jj) {
                const uint32_t i = cand[jj];
                if (!root || i == t) continue;
                const uint32_t oi = order[i];
                if (target[oi] != oj) continue;
                const double mi = spos[i].w;
                m += mi;
                for (int k = 0; k < 3; ++k) {
                    const double v = vx(oi, i, k);
                    mom[k] += mi * v;
                    com[k] += mi * px(oi, i, k);
                    ke += 0.5 * mi * v * v;
                }
                mu += mi * (double)svel[i].w;
                absorbed += 1;
            }
        }, sinst ? iroot[sinst[grp.x]] : 0u);
        if (root && absorbed > 0) {
            double v[3], x[3], v2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                v[k] = mom[k] / m;
                x[k] = com[k] / m;
                v2 += v[k] * v[k];
            }
            const double u_new = (mu + fmax(ke - 0.5 * m * v2, 0.0)) / m;  // destroyed relative KE -> heat
            const float h_new = hj * powf((float)(m / (double)me.w), 1.0f / 3.0f);
            posh[oj] = make_float4((float)x[0], (float)x[1], (float)x[2], h_new);
            velu[oj] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)u_new);
            mass[oj] = (float)m;
            if (pos64) {
                for (int k = 0; k < 3; ++k) {
                    pos64[3 * oj + k] = x[k];
                    vel64[3 * oj + k] = v[k];
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_merge_retire(uint32_t n, const uint32_t* __restrict__ target, float4* velu,
                                                      float* mass, double* vel64) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = target[i];
    if (t == kNoMerge || target[t] != kNoMerge) return;  // absorber was not a root: next pass
    mass[i] = 0.f;
    float4 v = velu[i];
    velu[i] = make_float4(0.f, 0.f, 0.f, v.w);
    if (vel64) vel64[3 * i] = vel64[3 * i + 1] = vel64[3 * i + 2] = 0.0;
}

void launch_merge(const Dev& d, const StepParams& p, hipStream_t s) {
    const uint32_t waves = std::min<uint32_t>(kWalkWaves, d.n);
    hipLaunchKernelGGL(k_merge_pick, dim3(waves), dim3(kWave), 0, s, d.spos, d.sh, d.svel, d.smat, d.scs, d.n, d.K,
                       d.L, d.nS, d.nrange, (uint32_t)p.sph_expand, d.vals, d.groups, d.scal, d.sinst(),
                       d.iroot, d.mtarget);
    hipLaunchKernelGGL(k_merge_apply, dim3(waves), dim3(kWave), 0, s, d.spos, d.sh, d.svel, d.n, d.K, d.L, d.nS,
                       d.nrange, (uint32_t)p.sph_expand, d.vals, d.groups, d.scal, d.sinst(), d.iroot,
                       d.mtarget, d.posh, d.velu, d.mass, d.pos64, d.vel64);
    hipLaunchKernelGGL(k_merge_retire, dim3((d.n + 255) / 256), dim3(256), 0, s, d.n, d.mtarget, d.velu, d.mass,
                       d.vel64);
}

}  // namespace isph
