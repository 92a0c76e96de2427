// Barnes-Hut gravity over the LBVH: wave-cooperative group traversal, plus the exact direct sum.
//
// Reference: bh_gravity.wgsl:337-394 (cs_gravity_bh: per-particle stack walk, robust MAC
// dist >= |bmax - bmin|/theta + |centre - com|, leaf buckets summed directly) and :313-328 / sph_step.wgsl:
// 235-241 (direct sum). Here one wave64 walks the tree for one walk group at once: a compact LBVH subtree
// of up to 64 x P particles (P per lane), taken from per-XCD work queues. A node is accepted only when the
// MAC holds for every point of the group's bounding box (unsoftened distance from the node COM to the box),
// so each particle sees a node set at least as refined as the reference's per-particle walk. Accepted
// cells and leaf particles are appended to LDS lists with ballot/mbcnt compaction and evaluated 64 at a
// time with LDS broadcast reads (P interactions per read).
// The traversal stack lives in LDS; up to 64 nodes are popped per round (one per lane), which leaves at
// most ~64 pending nodes per tree level (unrestricted, the deepest stack at N = 1M reaches ~540 entries).
// The pop width narrows towards depth-first near capacity, which bounds the stack by capacity - margin +
// tree depth. With a 512-entry stack the monopole walk's LDS is 4 KB per wave, so registers set its resident
// waves (8 per SIMD at 64 VGPRs); the quadrupole walk's extra lists take 9.5 KB, which caps it at 4.
#include "state.h"

namespace isph {

#ifndef ISPH_GRAVITY_STACK
#define ISPH_GRAVITY_STACK 512
#endif
constexpr int kStackCap = ISPH_GRAVITY_STACK;
constexpr int kStackMargin = 128;  // > maximum Karras depth (30 code bits + 32 index tie-break bits)
constexpr int kBuf = 128;

// Lane slots of a one-particle-per-lane monopole walk. The group's g particles and q strided chunks of each
// source batch (chunk c = batch entries c, c + q, c + 2q, ...) make q g (particle, chunk) slots: slot s is
// (particle s % g, chunk s / g), and lane l holds slots l, l + 64, ... (R = ceil(q g / 64) rounds). A
// batch of 64 sources then costs ceil(64 / q) R steps instead of 64. split_shape picks the q with the
// fewest steps among R <= rounds (q = 1: one slot per lane, the particle itself).
__device__ __forceinline__ void split_shape(uint32_t g, int rounds, int& q, int& R) {
    q = 1;
    R = 1;
    int best = kWave;
    for (int qq = 2; qq <= 16; ++qq) {
        const int rr = (int)((qq * g + kWave - 1) / kWave);
        const int cost = (kWave + qq - 1) / qq * rr;
        if (rr <= rounds && cost < best) {
            best = cost;
            q = qq;
            R = rr;
        }
    }
}

// Monopole walks put leaf particles into the cell list: with softening, a particle's own term has d = 0 and
// adds exactly zero, so the reference's s != t test (bh_gravity.wgsl:362) is only needed without softening
// (SELF), where the list also carries each entry's particle index. One list keeps the workgroup's LDS at
// ~4 KB, so the resident waves per SIMD are set by registers alone.
template <int P, bool QUAD, bool COMP, bool SELF>
__global__ __launch_bounds__(64, (P == 1 && !QUAD) ? 8 : 6) void k_gravity_tree(const float4* __restrict__ spos,
                                                                                uint32_t n, uint32_t K, uint32_t L, const uint4* __restrict__ nG,
                                                     const float4* __restrict__ nQ0, const float2* __restrict__ nQ1,
                                                     float soft2, float4* __restrict__ sgrav,
                                                     unsigned long long* counters, const uint2* __restrict__ groups,
                                                     uint32_t* scal, uint32_t count_slot,
                                                     const uint32_t* __restrict__ sinst,
                                                     const uint32_t* __restrict__ iroot) {
    constexpr bool UNI = !QUAD;
    constexpr bool SPLIT = P == 1 && UNI;          // lane slots (split_shape)
    constexpr int NR = SPLIT ? ((COMP || SELF) ? 2 : 3) : 1;  // slot rounds held per lane (no spills)
    constexpr int NA = SPLIT ? NR : P;              // accumulators per lane: one per slot or per particle
    static_assert(2 * NR * kWave <= kStackCap, "slot partials are reduced through the stack's LDS");
    __shared__ uint32_t stk[kStackCap];
    __shared__ float4 cA[kBuf];
    __shared__ float4 cQ0[QUAD ? kBuf : 1];
    __shared__ float2 cQ1[QUAD ? kBuf : 1];
    __shared__ uint32_t cI[(UNI && SELF) ? kBuf : 1];
    __shared__ float4 pA[UNI ? 1 : kBuf];
    __shared__ uint32_t pI[UNI ? 1 : kBuf];

    const int lane = threadIdx.x;
    const uint32_t leaf0 = L - 1;
    unsigned long long tot_cell = 0, tot_part = 0, rounds = 0, popped = 0;
    int max_sp = 1;
    WalkQueue queue(scal + kQueueGrav, scal[count_slot], 1);
    uint32_t gi;
    while (queue.next(gi)) {
        const uint2 grp = groups[gi];
        // Lane l holds particle l (+ 64 k for P = 2); with SPLIT its slot r takes a particle (read from that
        // particle's lane by a shuffle) and a batch chunk. Slots past q g are idle and take chunk 0; their
        // sums are never read.
        int q = 1, R = 1;
        if (SPLIT) split_shape(grp.y, NR, q, R);
        uint32_t slot_pc[NR];  // particle << 8 | chunk
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const uint32_t s = (uint32_t)(r * kWave + lane);
            const bool used = s < (uint32_t)q * grp.y;
            slot_pc[r] = SPLIT ? (s % grp.y) << 8 | (used ? s / grp.y : 0u) : (uint32_t)lane << 8;
        }
        uint32_t t[P];
        bool valid[P];
        float4 me[P];
        float lx0 = INFINITY, ly0 = INFINITY, lz0 = INFINITY, lx1 = -INFINITY, ly1 = -INFINITY, lz1 = -INFINITY;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const uint32_t slot = (uint32_t)lane + (uint32_t)k * kWave;
            t[k] = grp.x + slot;
            valid[k] = slot < grp.y;
            me[k] = spos[valid[k] ? t[k] : grp.x];
            lx0 = fminf(lx0, me[k].x); ly0 = fminf(ly0, me[k].y); lz0 = fminf(lz0, me[k].z);
            lx1 = fmaxf(lx1, me[k].x); ly1 = fmaxf(ly1, me[k].y); lz1 = fmaxf(lz1, me[k].z);
        }
        const float gx0 = uniform_f(wave_min(lx0)), gy0 = uniform_f(wave_min(ly0)), gz0 = uniform_f(wave_min(lz0));
        const float gx1 = uniform_f(wave_max(lx1)), gy1 = uniform_f(wave_max(ly1)), gz1 = uniform_f(wave_max(lz1));

        float ax[NA], ay[NA], az[NA];
        KahanF kx[NA], ky[NA], kz[NA];
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            ax[k] = ay[k] = az[k] = 0.f;
            kx[k] = KahanF{0.f, 0.f};
            ky[k] = KahanF{0.f, 0.f};
            kz[k] = KahanF{0.f, 0.f};
        }
        unsigned long long ncell = 0, npart = 0;

        auto add = [&](int k, float bx, float by, float bz) {
            if (COMP) {
                kx[k].add(bx);
                ky[k].add(by);
                kz[k].add(bz);
            } else {
                ax[k] += bx;
                ay[k] += by;
                az[k] += bz;
            }
        };
        // SPLIT: slot r walks its chunk's entries cr[r], cr[r] + q, ... of the batch; every chunk has cnt / q
        // whole steps, and the last step exists for the chunks below cnt % q only.
        auto flush_split = [&](int cnt) {
            const int full = cnt / q;
            const bool tail = full * q < cnt;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                if (r >= R) break;
                const int pr = (int)(slot_pc[r] >> 8);
                const float mx = __shfl(me[0].x, pr), my = __shfl(me[0].y, pr), mz = __shfl(me[0].z, pr);
                const uint32_t tt = grp.x + (uint32_t)pr;
                float bx = 0.f, by = 0.f, bz = 0.f;
                auto source = [&](int j, bool in) {
                    const float4 c = cA[j];
                    float dx = c.x - mx, dy = c.y - my, dz = c.z - mz;
                    float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
                    float inv = __builtin_amdgcn_rsqf(r2);
                    float w = c.w * inv * (inv * inv);
                    if (SELF) w = (cI[j] == tt) ? 0.f : w;  // s != t (bh_gravity.wgsl:362)
                    if (!in) w = 0.f;
                    bx += dx * w;
                    by += dy * w;
                    bz += dz * w;
                };
                int j = (int)(slot_pc[r] & 0xffu);
                for (int k = 0; k < full; ++k, j += q) source(j, true);
                if (tail) source(j < cnt ? j : 0, j < cnt);
                add(r, bx, by, bz);
            }
        };
        auto flush_cells = [&](int cnt) {
            if (SPLIT) {
                flush_split(cnt);
                return;
            }
            float bx[P], by[P], bz[P];
#pragma unroll
            for (int k = 0; k < P; ++k) bx[k] = by[k] = bz[k] = 0.f;
            auto source = [&](int j) {
                const float4 c = cA[j];
                const uint32_t idx = (UNI && SELF) ? cI[j] : 0u;
                float4 q0;
                float2 q1;
                if (QUAD) {
                    q0 = cQ0[j];
                    q1 = cQ1[j];
                }
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    float dx = c.x - me[k].x, dy = c.y - me[k].y, dz = c.z - me[k].z;
                    float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
                    float inv = __builtin_amdgcn_rsqf(r2);
                    float inv2 = inv * inv;
                    float w = c.w * inv * inv2;
                    if (QUAD) {
                        // g = G M / r^3 [d (1 + 2.5 (d.Qhat.d) / r^4) - Qhat.d / r^2], Qhat = Q / M; every
                        // intermediate stays within f32 range in SI units.
                        float qx = q0.x * dx + q0.w * dy + q1.x * dz;
                        float qy = q0.w * dx + q0.y * dy + q1.y * dz;
                        float qz = q1.x * dx + q1.y * dy + q0.z * dz;
                        float tq = (qx * dx + qy * dy + qz * dz) * inv2 * inv2;
                        float s = 1.0f + 2.5f * tq;
                        bx[k] += w * (dx * s - qx * inv2);
                        by[k] += w * (dy * s - qy * inv2);
                        bz[k] += w * (dz * s - qz * inv2);
                    } else {
                        if (UNI && SELF) w = (idx == t[k]) ? 0.f : w;  // s != t (bh_gravity.wgsl:362)
                        bx[k] += dx * w;
                        by[k] += dy * w;
                        bz[k] += dz * w;
                    }
                }
            };
            for (int j = 0; j < cnt; ++j) source(j);
#pragma unroll
            for (int k = 0; k < P; ++k) add(k, bx[k], by[k], bz[k]);
        };
        auto flush_parts = [&](int cnt) {
            float bx[P], by[P], bz[P];
#pragma unroll
            for (int k = 0; k < P; ++k) bx[k] = by[k] = bz[k] = 0.f;
            for (int j = 0; j < cnt; ++j) {
                const float4 c = pA[j];
                const uint32_t idx = pI[j];
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    float dx = c.x - me[k].x, dy = c.y - me[k].y, dz = c.z - me[k].z;
                    float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
                    float inv = __builtin_amdgcn_rsqf(r2);
                    float w = (idx == t[k]) ? 0.f : c.w * inv * inv * inv;  // s != t (bh_gravity.wgsl:362)
                    bx[k] += dx * w;
                    by[k] += dy * w;
                    bz[k] += dz * w;
                }
            }
#pragma unroll
            for (int k = 0; k < P; ++k) add(k, bx[k], by[k], bz[k]);
        };
        // Drop the first 64 entries of a list after they were evaluated.
        auto shift_cells = [&](int cnt) {
            __syncthreads();
            if (lane + kWave < cnt) {
                cA[lane] = cA[lane + kWave];
                if (UNI && SELF) cI[lane] = cI[lane + kWave];
                if (QUAD) {
                    cQ0[lane] = cQ0[lane + kWave];
                    cQ1[lane] = cQ1[lane + kWave];
                }
            }
            __syncthreads();
        };
        auto shift_parts = [&](int cnt) {
            __syncthreads();
            if (lane + kWave < cnt) {
                pA[lane] = pA[lane + kWave];
                pI[lane] = pI[lane + kWave];
            }
            __syncthreads();
        };

        int sp = 1, nc = 0, np = 0;
        __syncthreads();
        // root (a single leaf when L == 1), or the group's instance subtree in an ensemble
        if (lane == 0) stk[0] = sinst ? iroot[sinst[grp.x]] : 0u;
        __syncthreads();
        while (sp > 0) {
            int k = min(kWave, sp);
            int room = kStackCap - kStackMargin - sp;
            if (room < k) k = max(1, room);
            sp -= k;
            rounds += 1;
            popped += k;
            const bool has = lane < k;
            const uint32_t node = has ? stk[sp + lane] : 0u;
            __syncthreads();
            const bool is_leaf = has && node >= leaf0;
            bool accept = false, open = false;
            float4 A = make_float4(0.f, 0.f, 0.f, 0.f);
            uint32_t lc = 0, rc = 0;
            // Packed 32-byte record: (com, mass) then (MAC radius, left, right). One-particle leaves are never
            // tested (bh_gravity.wgsl:356-369) and use the first half only: it is the particle itself. The MAC
            // uses the unsoftened distance from the COM to the group box, d_min <= |x - com| for every member,
            // so acceptance implies the reference's softened per-particle MAC for each of them, and
            // d_min >= |bmax - bmin| / theta with theta < 1 keeps group particles outside accepted nodes.
            if (has) {
                const uint4 r0 = nG[2 * node];
                A = make_float4(__uint_as_float(r0.x), __uint_as_float(r0.y), __uint_as_float(r0.z),
                                __uint_as_float(r0.w));
            }
            if (has && (!is_leaf || K > 1)) {
                const uint4 r1 = nG[2 * node + 1];
                const float rcrit = __uint_as_float(r1.x);
                lc = r1.y;
                rc = r1.z;
                float dx = fmaxf(fmaxf(gx0 - A.x, A.x - gx1), 0.f);
                float dy = fmaxf(fmaxf(gy0 - A.y, A.y - gy1), 0.f);
                float dz = fmaxf(fmaxf(gz0 - A.z, A.z - gz1), 0.f);
                float d2 = dx * dx + dy * dy + dz * dz;
                accept = d2 > 0.f && d2 >= rcrit * rcrit;
                open = !accept && !is_leaf;
            }
            const bool leaf = is_leaf && !accept;
            const uint64_t mo = __ballot(open);
            if (open) {
                int pos = sp + 2 * prefix_count(mo);
                stk[pos] = lc;
                stk[pos + 1] = rc;
            }
            sp += 2 * __popcll(mo);
            max_sp = max(max_sp, sp);
            const uint64_t ma = __ballot(accept);
            if (accept) {
                int pos = nc + prefix_count(ma);
                cA[pos] = A;
                if (UNI && SELF) cI[pos] = 0xffffffffu;
                if (QUAD) {
                    cQ0[pos] = nQ0[node];
                    cQ1[pos] = nQ1[node];
                }
            }
            nc += __popcll(ma);
            ncell += __popcll(ma);
            const uint64_t ml = __ballot(leaf);
            if (UNI && ml && nc >= kWave) {  // room for a full batch of leaf particles
                __syncthreads();
                flush_cells(kWave);
                shift_cells(nc);
                nc -= kWave;
            }
            if (ml) {
                const uint32_t c = leaf ? node - leaf0 : 0u;
                const uint32_t start = c * K;
                const uint32_t end = leaf ? min(start + K, n) : 0u;
                for (uint32_t kk = 0; kk < K; ++kk) {
                    const bool present = leaf && start + kk < end;
                    const uint64_t mp = __ballot(present);
                    if (!mp) break;
                    if (present) {
                        // Two explicit stores: a conditional on the float4 lvalues would select between a
                        // register value and a global address, and forces A into private memory.
                        if (UNI) {
                            int pos = nc + prefix_count(mp);
                            if (K == 1) cA[pos] = A;
                            else cA[pos] = spos[start + kk];
                            if (SELF) cI[pos] = start + kk;
                        } else {
                            int pos = np + prefix_count(mp);
                            if (K == 1) pA[pos] = A;
                            else pA[pos] = spos[start + kk];
                            pI[pos] = start + kk;
                        }
                    }
                    npart += __popcll(mp);
                    if (UNI) {
                        nc += __popcll(mp);
                        if (nc >= kWave) {
                            __syncthreads();
                            flush_cells(kWave);
                            shift_cells(nc);
                            nc -= kWave;
                        }
                    } else {
                        np += __popcll(mp);
                        if (np >= kWave) {
                            __syncthreads();
                            flush_parts(kWave);
                            shift_parts(np);
                            np -= kWave;
                        }
                    }
                }
            }
            __syncthreads();
            if (nc >= kWave) {
                flush_cells(kWave);
                shift_cells(nc);
                nc -= kWave;
            }
            if (np >= kWave) {
                flush_parts(kWave);
                shift_parts(np);
                np -= kWave;
            }
        }
        if (nc > 0) flush_cells(nc);
        if (np > 0) flush_parts(np);
        if (SPLIT && q > 1) {
            // Particle p's sum is its q slots c g + p, added in chunk order through the stack's LDS (slot s at
            // red[s], Kahan carries at red[NR 64 + s]); a compensated sum takes s - c.
            float* red = reinterpret_cast<float*>(stk);
            const uint32_t g = grp.y;
            auto reduce = [&](float (&a)[NA], KahanF (&kk)[NA]) {
                __syncthreads();
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    if (r >= R) break;
                    red[r * kWave + lane] = COMP ? kk[r].s : a[r];
                    if (COMP) red[NR * kWave + r * kWave + lane] = kk[r].c;
                }
                __syncthreads();
                if ((uint32_t)lane < g) {
                    for (int c = 1; c < q; ++c) {
                        const uint32_t s = (uint32_t)c * g + (uint32_t)lane;
                        if (COMP) {
                            kk[0].add(red[s]);
                            kk[0].add(-red[NR * kWave + s]);
                        } else {
                            a[0] += red[s];
                        }
                    }
                }
            };
            reduce(ax, kx);
            reduce(ay, ky);
            reduce(az, kz);
        }
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (COMP) {
                ax[k] = kx[k].s;
                ay[k] = ky[k].s;
                az[k] = kz[k].s;
            }
            if (valid[k]) sgrav[t[k]] = make_float4(kG * ax[k], kG * ay[k], kG * az[k], 0.f);
        }
        tot_cell += ncell * grp.y;
        tot_part += npart * grp.y;
    }
    if (lane == 0 && counters) {
        unsigned long long* c = counters + (size_t)blockIdx.x * kCounterSlots;  // this wave's slots
        c[QUAD ? 2 : 0] += tot_cell;
        c[1] += tot_part;
        c[5] += rounds;
        c[6] += popped;
        c[7] = max(c[7], (unsigned long long)max_sp);
    }
}

// Exact softened direct sum in f32 (bh_gravity.wgsl:313-328), tiled through LDS.
template <bool COMP>
__global__ __launch_bounds__(256) void k_gravity_direct(const float4* __restrict__ spos, uint32_t n, float soft2,
                                                        float4* __restrict__ sgrav) {
    __shared__ float4 tile[256];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float4 me = spos[i < n ? i : n - 1];
    float ax = 0.f, ay = 0.f, az = 0.f;
    KahanF kx{0.f, 0.f}, ky{0.f, 0.f}, kz{0.f, 0.f};
    for (uint32_t base = 0; base < n; base += 256) {
        uint32_t j = base + threadIdx.x;
        tile[threadIdx.x] = j < n ? spos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        uint32_t cnt = min(256u, n - base);
        float bx = 0.f, by = 0.f, bz = 0.f;
        for (uint32_t jj = 0; jj < cnt; ++jj) {
            float4 c = tile[jj];
            float dx = c.x - me.x, dy = c.y - me.y, dz = c.z - me.z;
            float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
            float inv = __builtin_amdgcn_rsqf(r2);
            float w = (base + jj == i) ? 0.f : c.w * inv * inv * inv;
            bx += dx * w;
            by += dy * w;
            bz += dz * w;
        }
        if (COMP) {
            kx.add(bx);
            ky.add(by);
            kz.add(bz);
        } else {
            ax += bx;
            ay += by;
            az += bz;
        }
        __syncthreads();
    }
    if (COMP) {
        ax = kx.s;
        ay = ky.s;
        az = kz.s;
    }
    if (i < n) sgrav[i] = make_float4(kG * ax, kG * ay, kG * az, 0.f);
}

// ---- the f64 reference form ---------------------------------------------------------------------------
// The reference form runs the discrete step of the mixed form (same tree, same walk decisions, same pair
// sets, same dt) with f64 coordinates in every pair term and f64 gravity, so that the difference between a
// form's run and the reference run is that form's rounding alone. Coordinates are the f32 copies the walks
// read plus f32 residuals of the f64 state (x64 = x32 + lo, |lo| <= half an f32 ulp, so x32 + lo carries
// the f64 value to ~2^-48 of |x|).

__global__ __launch_bounds__(256) void k_gather_lo(uint32_t n, const uint32_t* __restrict__ order,
                                                   const double* __restrict__ pos64, const double* __restrict__ vel64,
                                                   const float4* __restrict__ spos, const float4* __restrict__ svel,
                                                   float4* __restrict__ slo, float4* __restrict__ svlo) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t o = order[t];
    const float4 p = spos[t], v = svel[t];
    slo[t] = make_float4((float)(pos64[3 * o] - (double)p.x), (float)(pos64[3 * o + 1] - (double)p.y),
                         (float)(pos64[3 * o + 2] - (double)p.z), 0.f);
    svlo[t] = make_float4((float)(vel64[3 * o] - (double)v.x), (float)(vel64[3 * o + 1] - (double)v.y),
                          (float)(vel64[3 * o + 2] - (double)v.z), 0.f);
}

// Depth of every internal node (parent hops to the root). A Karras child's common (code, index) prefix is
// longer than its parent's, and the key has 30 + 32 bits, so depths stay below kMaxTreeDepth.
constexpr uint32_t kMaxTreeDepth = 64;
__global__ __launch_bounds__(256) void k_node_depth(uint32_t L, const uint32_t* __restrict__ parent,
                                                    uint8_t* __restrict__ ndepth) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 >= L) return;
    uint32_t d = 0, p = i;
    while (d < kMaxTreeDepth) {
        const uint32_t q = parent[p];
        if (q == kNoParent) break;
        p = q;
        ++d;
    }
    ndepth[i] = (uint8_t)d;
}

__device__ __forceinline__ double4 body64(const float4* __restrict__ spos, const float4* __restrict__ slo, uint32_t s) {
    const float4 p = spos[s], lo = slo[s];
    return make_double4((double)p.x + (double)lo.x, (double)p.y + (double)lo.y, (double)p.z + (double)lo.z,
                        (double)p.w);
}

// f64 COM and mass of the internal nodes at one depth, from their children (leaves: the particles, K = 1).
// Launched from the deepest level up, so a node's children are complete when it is combined.
__global__ __launch_bounds__(256) void k_com64_level(uint32_t L, uint32_t level, const uint8_t* __restrict__ ndepth,
                                                     const uint32_t* __restrict__ left,
                                                     const uint32_t* __restrict__ right,
                                                     const float4* __restrict__ spos, const float4* __restrict__ slo,
                                                     double4* __restrict__ nA64) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 >= L || ndepth[i] != level) return;
    const uint32_t leaf0 = L - 1;
    const uint32_t lc = left[i], rc = right[i];
    const double4 a = lc >= leaf0 ? body64(spos, slo, lc - leaf0) : nA64[lc];
    const double4 b = rc >= leaf0 ? body64(spos, slo, rc - leaf0) : nA64[rc];
    const double m = a.w + b.w;
    double4 c;
    if (m > 0.0)
        c = make_double4((a.w * a.x + b.w * b.x) / m, (a.w * a.y + b.w * b.y) / m, (a.w * a.z + b.w * b.z) / m, m);
    else
        c = make_double4(0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z), 0.0);
    nA64[i] = c;
}

// The group walk of k_gravity_tree (one particle per lane, monopole, leaf size 1) with the same acceptance
// test on the same f32 node records, so the same nodes and leaf particles are taken; each accepted node
// contributes its f64 COM and mass and each leaf particle its f64 position, summed in f64 with IEEE sqrt
// and division. With softening a particle's own term has d = 0 and adds exactly zero; without, r^2 = 0
// terms are skipped.
__global__ __launch_bounds__(64) void k_gravity_tree_ref(const float4* __restrict__ spos,
                                                         const float4* __restrict__ slo, uint32_t n, uint32_t L,
                                                         const uint4* __restrict__ nG,
                                                         const double4* __restrict__ nA64, double soft2,
                                                         double4* __restrict__ sgrav64, float4* __restrict__ sgrav,
                                                         const uint2* __restrict__ groups, uint32_t* scal,
                                                         uint32_t count_slot) {
    __shared__ uint32_t stk[kStackCap];
    __shared__ double4 cA[kBuf];
    const int lane = threadIdx.x;
    const uint32_t leaf0 = L - 1;
    WalkQueue queue(scal + kQueueGrav, scal[count_slot], 1);
    uint32_t gi;
    while (queue.next(gi)) {
        const uint2 grp = groups[gi];
        const uint32_t t = grp.x + (uint32_t)lane;
        const bool valid = (uint32_t)lane < grp.y;
        const uint32_t tc = valid ? t : grp.x;
        const float4 me = spos[tc];
        const double4 me64 = body64(spos, slo, tc);
        const float gx0 = wave_min(me.x), gy0 = wave_min(me.y), gz0 = wave_min(me.z);
        const float gx1 = wave_max(me.x), gy1 = wave_max(me.y), gz1 = wave_max(me.z);
        double ax = 0.0, ay = 0.0, az = 0.0;
        auto flush = [&](int cnt) {
            for (int j = 0; j < cnt; ++j) {
                const double4 c = cA[j];
                const double dx = c.x - me64.x, dy = c.y - me64.y, dz = c.z - me64.z;
                const double r2 = dx * dx + dy * dy + dz * dz + soft2;
                const double w = r2 > 0.0 ? c.w / (r2 * sqrt(r2)) : 0.0;
                ax += dx * w;
                ay += dy * w;
                az += dz * w;
            }
        };
        auto shift = [&](int cnt) {
            __syncthreads();
            if (lane + kWave < cnt) cA[lane] = cA[lane + kWave];
            __syncthreads();
        };
        int sp = 1, nc = 0;
        __syncthreads();
        if (lane == 0) stk[0] = 0u;
        __syncthreads();
        while (sp > 0) {
            int k = min(kWave, sp);
            int room = kStackCap - kStackMargin - sp;
            if (room < k) k = max(1, room);
            sp -= k;
            const bool has = lane < k;
            const uint32_t node = has ? stk[sp + lane] : 0u;
            __syncthreads();
            const bool is_leaf = has && node >= leaf0;
            bool accept = false, open = false;
            uint32_t lc = 0, rc = 0;
            if (has && !is_leaf) {  // the acceptance test of k_gravity_tree, on the same f32 records
                const uint4 r0 = nG[2 * node];
                const uint4 r1 = nG[2 * node + 1];
                const float Ax = __uint_as_float(r0.x), Ay = __uint_as_float(r0.y), Az = __uint_as_float(r0.z);
                const float rcrit = __uint_as_float(r1.x);
                lc = r1.y;
                rc = r1.z;
                float dx = fmaxf(fmaxf(gx0 - Ax, Ax - gx1), 0.f);
                float dy = fmaxf(fmaxf(gy0 - Ay, Ay - gy1), 0.f);
                float dz = fmaxf(fmaxf(gz0 - Az, Az - gz1), 0.f);
                float d2 = dx * dx + dy * dy + dz * dz;
                accept = d2 > 0.f && d2 >= rcrit * rcrit;
                open = !accept;
            }
            const uint64_t mo = __ballot(open);
            if (open) {
                const int pos = sp + 2 * prefix_count(mo);
                stk[pos] = lc;
                stk[pos + 1] = rc;
            }
            sp += 2 * __popcll(mo);
            const uint64_t ma = __ballot(accept);
            if (accept) cA[nc + prefix_count(ma)] = nA64[node];
            nc += __popcll(ma);
            const uint64_t ml = __ballot(is_leaf);
            if (ml && nc >= kWave) {
                __syncthreads();
                flush(kWave);
                shift(nc);
                nc -= kWave;
            }
            if (ml) {
                if (is_leaf) cA[nc + prefix_count(ml)] = body64(spos, slo, node - leaf0);
                nc += __popcll(ml);
            }
            __syncthreads();
            if (nc >= kWave) {
                flush(kWave);
                shift(nc);
                nc -= kWave;
            }
        }
        if (nc > 0) flush(nc);
        if (valid) {
            sgrav64[t] = make_double4(kG64 * ax, kG64 * ay, kG64 * az, 0.0);
            sgrav[t] = make_float4((float)(kG64 * ax), (float)(kG64 * ay), (float)(kG64 * az), 0.f);
        }
    }
}

// The exact softened sum in f64 (IEEE sqrt and division), tiled through LDS.
__global__ __launch_bounds__(256) void k_gravity_direct_ref(const float4* __restrict__ spos,
                                                            const float4* __restrict__ slo, uint32_t n, double soft2,
                                                            double4* __restrict__ sgrav64, float4* __restrict__ sgrav) {
    __shared__ double4 tile[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double4 me = body64(spos, slo, i < n ? i : n - 1);
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (uint32_t base = 0; base < n; base += 256) {
        const uint32_t j = base + threadIdx.x;
        tile[threadIdx.x] = j < n ? body64(spos, slo, j) : make_double4(0.0, 0.0, 0.0, 0.0);
        __syncthreads();
        const uint32_t cnt = min(256u, n - base);
        for (uint32_t jj = 0; jj < cnt; ++jj) {
            const double4 c = tile[jj];
            const double dx = c.x - me.x, dy = c.y - me.y, dz = c.z - me.z;
            const double r2 = dx * dx + dy * dy + dz * dz + soft2;
            const double w = (base + jj == i || !(r2 > 0.0)) ? 0.0 : c.w / (r2 * sqrt(r2));
            ax += dx * w;
            ay += dy * w;
            az += dz * w;
        }
        __syncthreads();
    }
    if (i < n) {
        sgrav64[i] = make_double4(kG64 * ax, kG64 * ay, kG64 * az, 0.0);
        sgrav[i] = make_float4((float)(kG64 * ax), (float)(kG64 * ay), (float)(kG64 * az), 0.f);
    }
}

__global__ void k_fill4(float4* p, uint32_t n, float4 v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_scatter_gravity(uint32_t n, const uint32_t* __restrict__ order, const float4* __restrict__ sgrav,
                                  float4* __restrict__ acc) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) acc[order[t]] = sgrav[t];
}

void launch_gravity_tree(const Dev& d, const StepParams& p, hipStream_t s) {
    uint32_t waves = std::min<uint32_t>(kWalkWaves, d.n);
    // One particle per lane walks the SPH group list (groups of <= 64); two per lane use the gravity list.
    const bool own = p.grav_ppl == 2;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(waves), dim3(kWave), 0, s, d.spos, d.n, d.K, d.L, d.nG, d.nQ0, d.nQ1, p.soft2,
                           d.sgrav, d.counters, own ? d.ggroups : d.groups, d.scal,
                           (uint32_t)(own ? kGroupsGrav : kGroupsSph), d.sinst(), d.iroot);
    };
    // the monopole walks exclude a particle's own term explicitly only without softening
    const bool self = !p.quad && !(p.soft2 > 0.0f);
    const int sel = (p.grav_ppl == 2 ? 8 : 0) | (p.quad ? 4 : 0) | (p.mixed ? 2 : 0) | (self ? 1 : 0);
    switch (sel) {
        case 0: go(k_gravity_tree<1, false, false, false>); break;
        case 1: go(k_gravity_tree<1, false, false, true>); break;
        case 2: go(k_gravity_tree<1, false, true, false>); break;
        case 3: go(k_gravity_tree<1, false, true, true>); break;
        case 4: go(k_gravity_tree<1, true, false, false>); break;
        case 6: go(k_gravity_tree<1, true, true, false>); break;
        case 8: go(k_gravity_tree<2, false, false, false>); break;
        case 9: go(k_gravity_tree<2, false, false, true>); break;
        case 10: go(k_gravity_tree<2, false, true, false>); break;
        case 11: go(k_gravity_tree<2, false, true, true>); break;
        case 12: go(k_gravity_tree<2, true, false, false>); break;
        default: go(k_gravity_tree<2, true, true, false>); break;
    }
}
void launch_gravity_direct(const Dev& d, const StepParams& p, hipStream_t s) {
    if (p.mixed)
        hipLaunchKernelGGL(k_gravity_direct<true>, dim3((d.n + 255) / 256), dim3(256), 0, s, d.spos, d.n, p.soft2,
                           d.sgrav);
    else
        hipLaunchKernelGGL(k_gravity_direct<false>, dim3((d.n + 255) / 256), dim3(256), 0, s, d.spos, d.n, p.soft2,
                           d.sgrav);
}
void launch_gravity_direct_raw(const float4* bodies, uint32_t n, float soft2, float4* out, bool comp, hipStream_t s) {
    if (comp)
        hipLaunchKernelGGL(k_gravity_direct<true>, dim3((n + 255) / 256), dim3(256), 0, s, bodies, n, soft2, out);
    else
        hipLaunchKernelGGL(k_gravity_direct<false>, dim3((n + 255) / 256), dim3(256), 0, s, bodies, n, soft2, out);
}
void launch_gather_lo(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_lo, dim3((d.n + 255) / 256), dim3(256), 0, s, d.n, d.vals, d.pos64, d.vel64, d.spos,
                       d.svel, d.slo, d.svlo);
}
void launch_com64(const Dev& d, hipStream_t s) {
    if (d.L < 2) return;
    const uint32_t blocks = (d.L - 1 + 255) / 256;
    hipLaunchKernelGGL(k_node_depth, dim3(blocks), dim3(256), 0, s, d.L, d.parent, d.ndepth);
    for (uint32_t level = kMaxTreeDepth; level-- > 0;)
        hipLaunchKernelGGL(k_com64_level, dim3(blocks), dim3(256), 0, s, d.L, level, d.ndepth, d.left, d.right,
                           d.spos, d.slo, d.nA64);
}
void launch_gravity_tree_ref(const Dev& d, const StepParams& p, hipStream_t s) {
    const uint32_t waves = std::min<uint32_t>(kWalkWaves, d.n);
    hipLaunchKernelGGL(k_gravity_tree_ref, dim3(waves), dim3(kWave), 0, s, d.spos, d.slo, d.n, d.L, d.nG, d.nA64,
                       p.soft2_64, d.sgrav64, d.sgrav, d.groups, d.scal, (uint32_t)kGroupsSph);
}
void launch_gravity_direct_ref(const Dev& d, const StepParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_gravity_direct_ref, dim3((d.n + 255) / 256), dim3(256), 0, s, d.spos, d.slo, d.n, p.soft2_64,
                       d.sgrav64, d.sgrav);
}
__global__ void k_fill_d4(double4* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}
void launch_gravity_zero_ref(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_fill4, dim3((d.n + 255) / 256), dim3(256), 0, s, d.sgrav, d.n, make_float4(0.f, 0.f, 0.f, 0.f));
    hipLaunchKernelGGL(k_fill_d4, dim3((d.n + 255) / 256), dim3(256), 0, s, d.sgrav64, d.n);
}
void launch_gravity_zero(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_fill4, dim3((d.n + 255) / 256), dim3(256), 0, s, d.sgrav, d.n, make_float4(0.f, 0.f, 0.f, 0.f));
}
void launch_scatter_gravity(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_gravity, dim3((d.n + 255) / 256), dim3(256), 0, s, d.n, d.vals, d.sgrav, d.acc);
}

}  // namespace isph
