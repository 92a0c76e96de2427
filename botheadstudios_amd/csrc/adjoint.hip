// Adjoint of the particle step: loss, kick / kick-drift, gravity (direct and tree with the forward's
// acceptance decisions), SPH force + du/dt, per-particle Tillotson EOS (forward-mode duals for the local
// Jacobian) and SPH density, in the mixed form's arithmetic.
//
// Forward references: sph_step.wgsl:181-333 and hydrostatic.rs:338-465 (the differentiated step),
// bh_gravity.wgsl:337-394 (the walk whose interactions are replayed), eos.rs:52-96 (Tillotson).
// Pair loops are in gather form (each particle accumulates the adjoints of its own inputs from every pair
// it belongs to, using its neighbours' output adjoints), so the SPH adjoints need no atomics. The gravity
// walk's transposed interactions (the source side of particle-particle terms, and the node COM / mass
// adjoints) go through f64 atomics; node adjoints are then pushed down the tree to the particles.
#include "adjoint.h"
#include "walk.h"

namespace isph {

namespace {

__device__ __forceinline__ void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }

__device__ __forceinline__ double sigmoid(double z) { return 1.0 / (1.0 + exp(-z)); }

constexpr int kRedThreads = 256;

// Block sum of NV doubles into red[0..NV) (one atomic per block and value).
template <int NV>
__device__ __forceinline__ void block_sum_to(double (&v)[NV], double* red) {
    __shared__ double sh[kRedThreads / kWave][NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum_f64(v[k]);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[w][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int q = 0; q < (int)(blockDim.x / kWave); ++q) s += sh[q][threadIdx.x];
        atomic_add_f64(&red[threadIdx.x], s);
    }
}

}  // namespace

// ---- loss --------------------------------------------------------------------------------------------
// red[0] M, [1..3] sum m x, [4..6] sum m v, [7] loss, [8..10] Xbar, [11..13] Vbar, [14] Mbar.
__global__ __launch_bounds__(kRedThreads) void k_loss_moments(uint32_t n, const double* __restrict__ x,
                                                              const double* __restrict__ v,
                                                              const float* __restrict__ mass, double* red) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double m = mass[i];
        acc[0] += m;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            acc[1 + k] += m * x[3 * i + k];
            acc[4 + k] += m * v[3 * i + k];
        }
    }
    block_sum_to<7>(acc, red);
}

struct LossTerm {
    double l, dl_dr, dl_dj, dl_dM, s12;
};
__device__ __forceinline__ LossTerm loss_term(double m, double r, double j, double M, const LossParams& lp) {
    LossTerm o{0, 0, 0, 0, 0};
    if (!(m > 0.0) || !(r > 0.0)) return o;
    const double jc = sqrt(kG64 * M * r);
    const double s1 = sigmoid(lp.k_r * (r / lp.radius - lp.r_cut));
    const double s2 = sigmoid(lp.k_j * (j / jc - 1.0));
    const double ds1 = s1 * (1.0 - s1), ds2 = s2 * (1.0 - s2);
    o.s12 = s1 * s2;
    o.l = m * s1 * s2;
    o.dl_dr = m * (ds1 * (lp.k_r / lp.radius) * s2 - s1 * ds2 * lp.k_j * j / (2.0 * jc * r));
    o.dl_dj = m * s1 * ds2 * lp.k_j / jc;
    o.dl_dM = -m * s1 * ds2 * lp.k_j * j / (2.0 * jc * M);
    return o;
}

__global__ __launch_bounds__(kRedThreads) void k_loss_terms(uint32_t n, const double* __restrict__ x,
                                                            const double* __restrict__ v,
                                                            const float* __restrict__ mass, LossParams lp,
                                                            double scale, double* red, double* xb, double* vb,
                                                            double* ub, double* mb) {
    const double M = red[0];
    const double X[3] = {red[1] / M, red[2] / M, red[3] / M};
    const double V[3] = {red[4] / M, red[5] / M, red[6] / M};
    const double a[3] = {lp.ax, lp.ay, lp.az};
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // loss, Xbar, Vbar, Mbar
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double rel[3], w[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            rel[k] = x[3 * i + k] - X[k];
            w[k] = v[3 * i + k] - V[k];
        }
        const double r = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
        const double cr[3] = {rel[1] * w[2] - rel[2] * w[1], rel[2] * w[0] - rel[0] * w[2],
                              rel[0] * w[1] - rel[1] * w[0]};
        const double j = a[0] * cr[0] + a[1] * cr[1] + a[2] * cr[2];
        const LossTerm t = loss_term((double)mass[i], r, j, M, lp);
        // d j / d rel = w x a, d j / d w = a x rel
        const double wxa[3] = {w[1] * a[2] - w[2] * a[1], w[2] * a[0] - w[0] * a[2], w[0] * a[1] - w[1] * a[0]};
        const double axr[3] = {a[1] * rel[2] - a[2] * rel[1], a[2] * rel[0] - a[0] * rel[2],
                               a[0] * rel[1] - a[1] * rel[0]};
        const double inv_r = r > 0.0 ? 1.0 / r : 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double gr = t.dl_dr * rel[k] * inv_r + t.dl_dj * wxa[k];
            const double gw = t.dl_dj * axr[k];
            xb[3 * i + k] = scale * gr;
            vb[3 * i + k] = scale * gw;
            acc[1 + k] -= gr;
            acc[4 + k] -= gw;
        }
        ub[i] = 0.0;
        mb[i] = scale * t.s12;
        acc[0] += t.l;
        acc[7] += t.dl_dM;
    }
    block_sum_to<8>(acc, red + 7);
}

// Centre-of-mass terms: X = sum m x / M, V = sum m v / M and M = sum m feed back into every particle.
__global__ __launch_bounds__(256) void k_loss_finish(uint32_t n, const double* __restrict__ x,
                                                     const double* __restrict__ v, const float* __restrict__ mass,
                                                     const double* __restrict__ red, double scale, double* xb,
                                                     double* vb, double* mb) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double M = red[0];
    const double m = mass[i];
    double dm = red[14];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double Xk = red[1 + k] / M, Vk = red[4 + k] / M;
        const double Xb = red[8 + k], Vb = red[11 + k];
        xb[3 * i + k] += scale * (m / M) * Xb;
        vb[3 * i + k] += scale * (m / M) * Vb;
        dm += Xb * (x[3 * i + k] - Xk) / M + Vb * (v[3 * i + k] - Vk) / M;
    }
    mb[i] += scale * dm;
}

// ---- kicks (upload order; the forward is k_kick_drift / k_kick in their f64-state form) -----------------
// Forward: v1 = vh + a1 dt/2, u1 = max(uh + e1 dt/2, 0).
__global__ __launch_bounds__(256) void k_vjp_kick(uint32_t n, const float* __restrict__ mass,
                                                  const float* __restrict__ u_mid, const float4* __restrict__ acc,
                                                  float dt, const double* __restrict__ vb, double* ub,
                                                  double4* ab) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(mass[i] > 0.f)) {  // retired: not integrated, the adjoint passes through
        ab[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
    }
    const double hd = 0.5 * (double)dt;
    const float he = 0.5f * dt;
    const bool live_u = u_mid[i] + acc[i].w * he > 0.0f;
    const double u1b = ub[i];
    ab[i] = make_double4(vb[3 * i] * hd, vb[3 * i + 1] * hd, vb[3 * i + 2] * hd, live_u ? u1b * (double)he : 0.0);
    ub[i] = live_u ? u1b : 0.0;
}

// Forward: vh = v + a0 dt/2, uh = max(u + e0 dt/2, 0), x1 = x + vh dt.
__global__ __launch_bounds__(256) void k_vjp_kick_drift(uint32_t n, const float* __restrict__ mass,
                                                        const float* __restrict__ u_start,
                                                        const float4* __restrict__ acc, float dt,
                                                        const double* __restrict__ xb, double* vb, double* ub,
                                                        double4* ab) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(mass[i] > 0.f)) {
        ab[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
    }
    const double hd = 0.5 * (double)dt;
    const float he = 0.5f * dt;
    double w[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w[k] = vb[3 * i + k] + xb[3 * i + k] * (double)dt;
        vb[3 * i + k] = w[k];
    }
    const bool live_u = u_start[i] + acc[i].w * he > 0.0f;
    const double uhb = ub[i];
    ab[i] = make_double4(w[0] * hd, w[1] * hd, w[2] * hd, live_u ? uhb * (double)he : 0.0);
    ub[i] = live_u ? uhb : 0.0;
}

// ---- per-evaluation adjoint (sorted order) ---------------------------------------------------------
__global__ __launch_bounds__(256) void k_vjp_gather(uint32_t n, const uint32_t* __restrict__ order,
                                                    const float4* __restrict__ spos, const double4* __restrict__ ab,
                                                    double4* sa) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    sa[t] = spos[t].w > 0.f ? ab[order[t]] : make_double4(0.0, 0.0, 0.0, 0.0);
}

// Gravity, exact sum: a_i = G sum_j m_j d s, d = x_j - x_i, s = (|d|^2 + eps^2)^-1.5. With
// g_ij = G m_j s (abar_i - 3 (d.abar_i) d / (|d|^2 + eps^2)) (the source-side gradient),
// xbar_i = sum_j (g_ji - g_ij) and mbar_i = G sum_j s (x_i - x_j).abar_j. Pair arithmetic in f32 (as the
// forward), tile sums in f64.
constexpr int kDirTile = 256;
__global__ __launch_bounds__(kDirTile) void k_vjp_gravity_direct(uint32_t n, const float4* __restrict__ spos,
                                                                 const double4* __restrict__ sa, float soft2,
                                                                 double4* sxb) {
    __shared__ float4 tp[kDirTile];
    __shared__ float4 ta[kDirTile];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ic = i < n ? i : n - 1;
    const float4 me = spos[ic];
    const double4 a64 = sa[ic];
    const float ax = (float)a64.x, ay = (float)a64.y, az = (float)a64.z;
    double gx = 0, gy = 0, gz = 0, gm = 0;
    for (uint32_t base = 0; base < n; base += kDirTile) {
        uint32_t j = base + threadIdx.x;
        __syncthreads();
        if (j < n) {
            tp[threadIdx.x] = spos[j];
            double4 b = sa[j];
            ta[threadIdx.x] = make_float4((float)b.x, (float)b.y, (float)b.z, 0.f);
        } else {
            tp[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
            ta[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        float fx = 0.f, fy = 0.f, fz = 0.f, fm = 0.f;
        const uint32_t cnt = min((uint32_t)kDirTile, n - base);
        for (uint32_t k = 0; k < cnt; ++k) {
            if (base + k == i) continue;
            const float4 pj = tp[k];
            const float4 aj = ta[k];
            const float dx = pj.x - me.x, dy = pj.y - me.y, dz = pj.z - me.z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
            const float inv = __builtin_amdgcn_rsqf(r2);
            const float inv2 = inv * inv;
            const float s = inv * inv2;
            const float dai = dx * ax + dy * ay + dz * az;
            const float daj = dx * aj.x + dy * aj.y + dz * aj.z;
            const float ci = 3.0f * dai * inv2, cj = 3.0f * daj * inv2;
            // m_i (abar_j - 3 (d.abar_j) d / r2) - m_j (abar_i - 3 (d.abar_i) d / r2), scaled by s first: in SI
            // units a particle mass (~1e19 kg) times an acceleration adjoint leaves the f32 range over long
            // horizons, while s m stays small
            const float smi = s * me.w, smj = s * pj.w;
            fx += smi * (aj.x - cj * dx) - smj * (ax - ci * dx);
            fy += smi * (aj.y - cj * dy) - smj * (ay - ci * dy);
            fz += smi * (aj.z - cj * dz) - smj * (az - ci * dz);
            fm -= s * daj;
        }
        gx += fx;
        gy += fy;
        gz += fz;
        gm += fm;
    }
    if (i < n) {
        double4 o = sxb[i];
        o.x += kG64 * gx;
        o.y += kG64 * gy;
        o.z += kG64 * gz;
        o.w += kG64 * gm;
        sxb[i] = o;
    }
}

// Gravity, tree: replays the forward walk of each group (same groups, same MAC, same node records) and
// differentiates its interactions. Within a flush of up to 64 sources, lane l owns source l; at rotation k
// lane l evaluates the pair (source l, particle (l + k) mod 64) and keeps the source-side gradient in its
// registers. The particle-side sums travel: lane l holds particle (l + k) mod 64's partial sum at rotation
// k, and a wave rotate by one lane (DPP wave_rol, dst[l] = src[l + 1]) hands it on, so after 64 rotations
// every sum is home. Source sums leave through f64 atomics: node COM / mass adjoints (pushed down to
// particles afterwards) and the partner particles' position / mass adjoints.
constexpr int kGStack = 512;
constexpr int kGStackMargin = 128;
constexpr int kGBuf = 128;

__global__ __launch_bounds__(64) void k_vjp_gravity_tree(const float4* __restrict__ spos, uint32_t n, uint32_t K,
                                                         uint32_t L, const uint4* __restrict__ nG, float soft2,
                                                         const double4* __restrict__ sa,
                                                         const uint2* __restrict__ groups,
                                                         const uint32_t* __restrict__ scal, double4* nbar,
                                                         uint32_t nbar_stride,
                                                         double4* sxb) {
    __shared__ uint32_t stk[kGStack];
    __shared__ float4 cA[kGBuf];
    __shared__ uint32_t cN[kGBuf];
    __shared__ float4 pA[kGBuf];
    __shared__ uint32_t pI[kGBuf];
    __shared__ float4 gP[kWave];  // group particle positions, w = 1 if valid
    __shared__ float4 gB[kWave];  // group particle abar (f32)
    __shared__ uint32_t gI[kWave];  // group particle sorted index (0xffffffff when the lane is empty)
    const int lane = threadIdx.x;
    const uint32_t leaf0 = L - 1;
    const uint32_t ngroups = scal[kGroupsSph];
    for (uint32_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const uint2 grp = groups[gi];
        const uint32_t t = grp.x + lane;
        const bool valid = (uint32_t)lane < grp.y;
        const float4 me = spos[valid ? t : grp.x];
        const double4 a64 = valid ? sa[t] : make_double4(0.0, 0.0, 0.0, 0.0);
        const float gx0 = wave_min(me.x), gy0 = wave_min(me.y), gz0 = wave_min(me.z);
        const float gx1 = wave_max(me.x), gy1 = wave_max(me.y), gz1 = wave_max(me.z);
        __syncthreads();
        gP[lane] = make_float4(me.x, me.y, me.z, valid ? 1.f : 0.f);
        gB[lane] = make_float4((float)a64.x, (float)a64.y, (float)a64.z, 0.f);
        gI[lane] = valid ? t : 0xffffffffu;
        double xbx = 0, xby = 0, xbz = 0;

        // Rotation over the 64 group particles for sources held by lanes < cnt. `self_check`: the sources
        // are particles, which skip their own pair (cells never coincide with a group particle's index).
        auto rol1 = [](float v) { return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xF, 0xF, false)); };
        auto rotate = [&](int cnt, const float4 src, const uint32_t src_idx, bool self_check, float& sx, float& sy,
                          float& sz, float& sm) {
            float px = 0.f, py = 0.f, pz = 0.f;
            sx = sy = sz = sm = 0.f;
            for (int k = 0; k < kWave; ++k) {
                if (k > 0) {
                    px = rol1(px);
                    py = rol1(py);
                    pz = rol1(pz);
                }
                const int pidx = (lane + k) & (kWave - 1);
                const float4 xp = gP[pidx];
                const float4 bp = gB[pidx];
                float gx = 0.f, gy = 0.f, gz = 0.f;
                if (lane < cnt && xp.w > 0.f && (!self_check || gI[pidx] != src_idx)) {
                    const float dx = src.x - xp.x, dy = src.y - xp.y, dz = src.z - xp.z;
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    const float inv2 = inv * inv;
                    const float s = inv * inv2;
                    const float da = dx * bp.x + dy * bp.y + dz * bp.z;
                    const float c = 3.0f * da * inv2;
                    gx = src.w * s * (bp.x - c * dx);
                    gy = src.w * s * (bp.y - c * dy);
                    gz = src.w * s * (bp.z - c * dz);
                    sx += gx;
                    sy += gy;
                    sz += gz;
                    sm += s * da;
                }
                // the particle-side gradient of pair (lane, pidx) joins pidx's travelling sum
                px -= gx;
                py -= gy;
                pz -= gz;
            }
            xbx += rol1(px);
            xby += rol1(py);
            xbz += rol1(pz);
        };
        auto flush_cells = [&](int cnt) {
            __syncthreads();
            const float4 src = lane < cnt ? cA[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            const uint32_t node = lane < cnt ? cN[lane] : 0u;
            float sx, sy, sz, sm;
            rotate(cnt, src, 0xfffffffeu, false, sx, sy, sz, sm);
            if (lane < cnt) {
                double* nb = reinterpret_cast<double*>(
                    &nbar[(size_t)(blockIdx.x % kNbarSlots) * nbar_stride + node]);
                atomic_add_f64(nb + 0, kG64 * sx);
                atomic_add_f64(nb + 1, kG64 * sy);
                atomic_add_f64(nb + 2, kG64 * sz);
                atomic_add_f64(nb + 3, kG64 * sm);
            }
        };
        auto flush_parts = [&](int cnt) {
            __syncthreads();
            const float4 src = lane < cnt ? pA[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            const uint32_t idx = lane < cnt ? pI[lane] : 0xfffffffeu;
            float sx, sy, sz, sm;
            rotate(cnt, src, idx, true, sx, sy, sz, sm);
            if (lane < cnt) {
                double* pb = reinterpret_cast<double*>(&sxb[idx]);
                atomic_add_f64(pb + 0, kG64 * sx);
                atomic_add_f64(pb + 1, kG64 * sy);
                atomic_add_f64(pb + 2, kG64 * sz);
                atomic_add_f64(pb + 3, kG64 * sm);
            }
        };
        auto shift_cells = [&](int cnt) {
            __syncthreads();
            if (lane + kWave < cnt) {
                cA[lane] = cA[lane + kWave];
                cN[lane] = cN[lane + kWave];
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
        if (lane == 0) stk[0] = 0;
        __syncthreads();
        while (sp > 0) {
            int k = min(kWave, sp);
            int room = kGStack - kGStackMargin - sp;
            if (room < k) k = max(1, room);
            sp -= k;
            const bool has = lane < k;
            const uint32_t node = has ? stk[sp + lane] : 0u;
            __syncthreads();
            const bool is_leaf = has && node >= leaf0;
            bool accept = false, open = false;
            float4 A = make_float4(0.f, 0.f, 0.f, 0.f);
            uint32_t lc = 0, rc = 0;
            if (has) {
                const uint4 r0 = nG[2 * node];
                A = make_float4(__uint_as_float(r0.x), __uint_as_float(r0.y), __uint_as_float(r0.z),
                                __uint_as_float(r0.w));
            }
            if (has && (!is_leaf || K > 1)) {  // the forward's MAC (k_gravity_tree)
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
            const uint64_t ma = __ballot(accept);
            if (accept) {
                int pos = nc + prefix_count(ma);
                cA[pos] = A;
                cN[pos] = node;
            }
            nc += __popcll(ma);
            const uint64_t ml = __ballot(leaf);
            if (ml) {
                const uint32_t c = leaf ? node - leaf0 : 0u;
                const uint32_t start = c * K;
                const uint32_t end = leaf ? min(start + K, n) : 0u;
                for (uint32_t kk = 0; kk < K; ++kk) {
                    const bool present = leaf && start + kk < end;
                    const uint64_t mp = __ballot(present);
                    if (!mp) break;
                    if (present) {
                        int pos = np + prefix_count(mp);
                        if (K == 1) pA[pos] = A;
                        else pA[pos] = spos[start + kk];
                        pI[pos] = start + kk;
                    }
                    np += __popcll(mp);
                    if (np >= kWave) {
                        flush_parts(kWave);
                        shift_parts(np);
                        np -= kWave;
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
        if (valid) {
            double* pb = reinterpret_cast<double*>(&sxb[t]);
            atomic_add_f64(pb + 0, kG64 * xbx);
            atomic_add_f64(pb + 1, kG64 * xby);
            atomic_add_f64(pb + 2, kG64 * xbz);
        }
    }
}

// Folds the kNbarSlots partial node sums into slot 0.
__global__ __launch_bounds__(256) void k_vjp_nbar_reduce(uint32_t nodes, double4* nbar) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nodes) return;
    double4 s = nbar[t];
#pragma unroll
    for (uint32_t k = 1; k < kNbarSlots; ++k) {
        const double4 b = nbar[(size_t)k * nodes + t];
        s.x += b.x;
        s.y += b.y;
        s.z += b.z;
        s.w += b.w;
    }
    nbar[t] = s;
}

// Node adjoints to particles: COM_n = sum m_k x_k / M_n and M_n = sum m_k over the particles under n.
__global__ __launch_bounds__(256) void k_vjp_downsweep(uint32_t n, uint32_t K, uint32_t L,
                                                       const uint32_t* __restrict__ parent,
                                                       const float4* __restrict__ nA,
                                                       const double4* __restrict__ nbar,
                                                       const float4* __restrict__ spos, double4* sxb) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float4 p = spos[t];
    const double m = p.w;
    double bx = 0, by = 0, bz = 0, bm = 0;
    uint32_t node = (L - 1) + t / K;
    while (node != kNoParent) {
        const double4 b = nbar[node];
        if (b.x != 0.0 || b.y != 0.0 || b.z != 0.0 || b.w != 0.0) {
            const float4 A = nA[node];
            const double M = A.w;
            if (M > 0.0) {
                bx += m / M * b.x;
                by += m / M * b.y;
                bz += m / M * b.z;
                bm += b.w + (((double)p.x - A.x) * b.x + ((double)p.y - A.y) * b.y + ((double)p.z - A.z) * b.z) / M;
            }
        }
        node = parent[node];
    }
    double4 o = sxb[t];
    o.x += bx;
    o.y += by;
    o.z += bz;
    o.w += bm;
    sxb[t] = o;
}

// SPH force + du/dt adjoint (forward: k_force<true>, sph_step.wgsl:255-295 in f64 pair arithmetic). For a
// pair with d = x_i - x_j, v = v_i - v_j, F = W'(r)/r, C = q_i + q_j + Pi_ij, the pair adds
// -m_j C F d and +m_i C F d to a_i, a_j and m_j C F (v.d)/2, m_i C F (v.d)/2 to du_i/dt, du_j/dt; with
// A = -m_j abar_i + m_i abar_j and D = (m_j ebar_i + m_i ebar_j)/2 its loss is C F (A + D v).d.
__device__ __forceinline__ double spline_d2w(double q, double inv_h, double sig4) {
    // second derivative of W in r: sig4 = 8/(pi h^4); d/dr (sig4 (18 q^2 - 12 q)) etc.
    return q < 0.5 ? sig4 * inv_h * (36.0 * q - 12.0) : sig4 * inv_h * 12.0 * (1.0 - q);
}

__global__ __launch_bounds__(64) void k_vjp_force(const float4* __restrict__ spos, const float* __restrict__ sh,
                                                  const float4* __restrict__ svel, uint32_t n, uint32_t K,
                                                  uint32_t L, const SphNode* __restrict__ nS,
                                                  const uint2* __restrict__ nrange, uint32_t expand,
                                                  const double* __restrict__ srho64,
                                                  const double* __restrict__ spr64,
                                                  const double* __restrict__ scs64, float av_alpha, float av_beta,
                                                  const double4* __restrict__ sa, const uint2* __restrict__ groups,
                                                  const uint32_t* __restrict__ scal, double4* sxb, double4* svb,
                                                  double* srb, double* sqb, double* scb) {
    __shared__ uint32_t stk[kSphStack];
    __shared__ uint32_t cand[kSphBuf];
    __shared__ float4 cP[kWave];   // x, y, z, h/2
    __shared__ float4 cV[kWave];   // vx, vy, vz, m
    __shared__ double cR[kWave], cQ[kWave], cC[kWave];
    __shared__ double4 cAb[kWave];
    __shared__ uint32_t own[kWave];
    const int lane = threadIdx.x;
    own[lane] = kWave;
    const double alpha = av_alpha, beta = av_beta;
    const uint32_t ngroups = scal[kGroupsSph];
    for (uint32_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const uint2 grp = groups[gi];
        const uint32_t t = grp.x + lane;
        const bool valid = (uint32_t)lane < grp.y;
        const uint32_t tc = valid ? t : grp.x;
        __syncthreads();
        const float4 me = spos[tc];
        const float hi = sh[tc];
        const float qh = 0.5f * hi;  // half smoothing length, as the candidates carry it
        const float4 vi = svel[tc];
        const double rhoi = srho64[tc], qi = spr64[tc], ci = scs64[tc];
        const double4 abi = valid ? sa[tc] : make_double4(0.0, 0.0, 0.0, 0.0);
        const float gx0 = wave_min(me.x), gy0 = wave_min(me.y), gz0 = wave_min(me.z);
        const float gx1 = wave_max(me.x), gy1 = wave_max(me.y), gz1 = wave_max(me.z);
        const float hg = wave_max(hi);
        double xb[3] = {0, 0, 0}, vbar[3] = {0, 0, 0}, rb = 0, qb = 0, cb = 0, mb = 0;
        neighbour_walk(stk, cand, n, K, L, nS, expand, gx0, gy0, gz0, gx1, gy1, gz1, hg, nullptr,
                       [&](int cnt) {
            if (lane < cnt) {
                uint32_t j = cand[lane];
                const float4 p = spos[j];
                cP[lane] = make_float4(p.x, p.y, p.z, 0.5f * sh[j]);
                const float4 v = svel[j];
                cV[lane] = make_float4(v.x, v.y, v.z, p.w);
                cR[lane] = srho64[j];
                cQ[lane] = spr64[j];
                cC[lane] = scs64[j];
                cAb[lane] = sa[j];
                note_owner(own, j, grp.x, grp.y, lane);
            }
            __syncthreads();
            // the forward's pair set: its f32 radius test, coincident pairs skipped per hit
            uint64_t rest = valid ? pair_mask(cP, cnt, me, qh) & ~self_bit(own, cand, cnt, lane, t) : 0ull;
            while (__ballot(rest != 0ull)) {
                if (!rest) continue;
                const int jj = __builtin_ctzll(rest);
                rest &= rest - 1ull;
                const float4 pj = cP[jj];
                const double d[3] = {(double)me.x - (double)pj.x, (double)me.y - (double)pj.y,
                                     (double)me.z - (double)pj.z};
                const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const double hij = (double)qh + (double)pj.w;
                if (!(r2 > 1.0e-18)) continue;
                const float4 vj = cV[jj];
                const double v[3] = {(double)vi.x - (double)vj.x, (double)vi.y - (double)vj.y,
                                     (double)vi.z - (double)vj.z};
                const double inv_r = nr_rsq(r2);
                const double r = r2 * inv_r;
                const double inv_h = nr_rcp(hij);
                const double q = r * inv_h;
                const double ih2 = inv_h * inv_h;
                const double sig4 = (8.0 / 3.14159265358979323846) * (ih2 * ih2);
                const double dw = q < 0.5 ? sig4 * (18.0 * q * q - 12.0 * q) : sig4 * -6.0 * (1.0 - q) * (1.0 - q);
                const double F = dw * inv_r;
                const double dF = (spline_d2w(q, inv_h, sig4) - F) * inv_r;  // dF/dr
                const double vr = v[0] * d[0] + v[1] * d[1] + v[2] * d[2];
                const double rhoj = cR[jj], qj = cQ[jj], cj = cC[jj];
                const double mj = vj.w, mi = me.w;
                double Pi = 0.0, dPi_dmu = 0.0, mu = 0.0, rden = 0.0, rrs = 0.0;
                if (vr < 0.0) {
                    rden = nr_rcp(r2 + 0.01 * hij * hij);
                    rrs = nr_rcp(rhoi + rhoj);
                    mu = hij * vr * rden;
                    const double cbar = 0.5 * (ci + cj);
                    Pi = 2.0 * (-alpha * cbar * mu + beta * mu * mu) * rrs;
                    dPi_dmu = 2.0 * (-alpha * cbar + 2.0 * beta * mu) * rrs;
                }
                const double C = qi + qj + Pi;
                const double4 abj = cAb[jj];
                const double A[3] = {-mj * abi.x + mi * abj.x, -mj * abi.y + mi * abj.y, -mj * abi.z + mi * abj.z};
                const double D = 0.5 * (mj * abi.w + mi * abj.w);
                const double wv[3] = {A[0] + D * v[0], A[1] + D * v[1], A[2] + D * v[2]};
                const double wvd = wv[0] * d[0] + wv[1] * d[1] + wv[2] * d[2];
                const double S = F * wvd;  // d loss / d C
                // position: C (F' (d/r) (wv.d) + F wv) + S dPi/dd
                double gd[3], gv[3];
                const double cdf = C * dF * inv_r * wvd, cf = C * F, cfd = cf * D;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    gd[k] = cdf * d[k] + cf * wv[k];
                    gv[k] = cfd * d[k];
                }
                if (vr < 0.0) {
                    const double sdp = S * dPi_dmu;
                    const double dmu_dvr = hij * rden, dmu_dr2 = -mu * rden;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        gd[k] += sdp * (dmu_dvr * v[k] + dmu_dr2 * 2.0 * d[k]);
                        gv[k] += sdp * dmu_dvr * d[k];
                    }
                    cb += S * (-alpha * mu * rrs);
                    rb += S * (-Pi * rrs);
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    xb[k] += gd[k];
                    vbar[k] += gv[k];
                }
                qb += S;
                mb += cf * ((abj.x + 0.5 * abj.w * v[0]) * d[0] + (abj.y + 0.5 * abj.w * v[1]) * d[1] +
                            (abj.z + 0.5 * abj.w * v[2]) * d[2]);
            }
        });
        if (valid) {
            double4 o = sxb[t];
            o.x += xb[0];
            o.y += xb[1];
            o.z += xb[2];
            o.w += mb;
            sxb[t] = o;
            double4 w = svb[t];
            w.x += vbar[0];
            w.y += vbar[1];
            w.z += vbar[2];
            svb[t] = w;
            srb[t] += rb;
            sqb[t] = qb;
            scb[t] = cb;
        }
    }
}

// Per-particle EOS adjoint: q = P / rho^2 and c = sqrt(c^2) as functions of (rho, u, a, b, A, B) through
// the same central-difference sound speed as the forward, differentiated with forward-mode duals.
constexpr int kEosAdjThreads = 256;
__global__ __launch_bounds__(kEosAdjThreads) void k_vjp_eos(uint32_t n, const float4* __restrict__ svel,
                                                            const uint32_t* __restrict__ smat,
                                                            const Till* __restrict__ eos,
                                                            const double* __restrict__ srho64,
                                                            const double* __restrict__ sqb,
                                                            const double* __restrict__ scb, double* srb,
                                                            double4* svb, double* pbar) {
    __shared__ double pb[kMaxMaterials * 4];
    for (int k = threadIdx.x; k < kMaxMaterials * 4; k += blockDim.x) pb[k] = 0.0;
    __syncthreads();
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const double qbar = sqb[t], cbar = scb[t];
        if (qbar == 0.0 && cbar == 0.0) continue;
        const uint32_t m = smat[t];
        const Till& T = eos[m];
        using D6 = Dual<6>;
        TillT<D6> td;
        td.rho0 = dual_const<6>(T.rho0);
        td.a = dual_var<6>(T.a, 2);
        td.b = dual_var<6>(T.b, 3);
        td.cap_a = dual_var<6>(T.cap_a, 4);
        td.cap_b = dual_var<6>(T.cap_b, 5);
        td.e0 = dual_const<6>(T.e0);
        td.e_iv = dual_const<6>(T.e_iv);
        td.e_cv = dual_const<6>(T.e_cv);
        td.alpha = dual_const<6>(T.alpha);
        td.beta = dual_const<6>(T.beta);
        const D6 rho = dual_var<6>(srho64[t], 0);
        const D6 u = dual_var<6>((double)svel[t].w, 1);
        D6 p, c2;
        till_p_c2_t(td, rho, u, p, c2);
        const D6 q = p / (rho * rho);
        const D6 c = gsqrt(c2);
        srb[t] += qbar * q.d[0] + cbar * c.d[0];
        double4 w = svb[t];
        w.w += qbar * q.d[1] + cbar * c.d[1];
        svb[t] = w;
#pragma unroll
        for (int k = 0; k < 4; ++k) atomicAdd(&pb[m * 4 + k], qbar * q.d[2 + k] + cbar * c.d[2 + k]);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxMaterials * 4; k += blockDim.x)
        if (pb[k] != 0.0) atomic_add_f64(&pbar[k], pb[k]);
}

// SPH density adjoint (forward: k_density<true>): rho_i = m_i W(0, h_i) + sum_j m_j W(r_ij, hbar_ij) over
// live particles. xbar_i += sum_j (rbar_i m_j + rbar_j m_i) W'(r) d / r, mbar_i += rbar_i W(0, h_i) +
// sum_j rbar_j W(r). Retired particles keep their density: their rbar does not propagate.
__global__ __launch_bounds__(64) void k_vjp_density(const float4* __restrict__ spos, const float* __restrict__ sh,
                                                    uint32_t n, uint32_t K, uint32_t L,
                                                    const SphNode* __restrict__ nS, const uint2* __restrict__ nrange,
                                                    uint32_t expand, const double* __restrict__ srb,
                                                    const uint2* __restrict__ groups,
                                                    const uint32_t* __restrict__ scal, double4* sxb) {
    __shared__ uint32_t stk[kSphStack];
    __shared__ uint32_t cand[kSphBuf];
    __shared__ float4 cP[kWave];  // x, y, z, h/2
    __shared__ float cM[kWave];
    __shared__ double cRb[kWave];
    __shared__ uint32_t own[kWave];
    const int lane = threadIdx.x;
    own[lane] = kWave;
    const uint32_t ngroups = scal[kGroupsSph];
    for (uint32_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const uint2 grp = groups[gi];
        const uint32_t t = grp.x + lane;
        const bool valid = (uint32_t)lane < grp.y;
        const uint32_t tc = valid ? t : grp.x;
        __syncthreads();
        const float4 me = spos[tc];
        const float hi = sh[tc];
        const float qh = 0.5f * hi;  // half smoothing length, as the candidates carry it
        const bool live = valid && me.w > 0.f;
        const double rbi = live ? srb[tc] : 0.0;
        const float gx0 = wave_min(me.x), gy0 = wave_min(me.y), gz0 = wave_min(me.z);
        const float gx1 = wave_max(me.x), gy1 = wave_max(me.y), gz1 = wave_max(me.z);
        const float hg = wave_max(hi);
        double xb[3] = {0, 0, 0}, mb = 0;
        neighbour_walk(stk, cand, n, K, L, nS, expand, gx0, gy0, gz0, gx1, gy1, gz1, hg, nullptr,
                       [&](int cnt) {
            if (lane < cnt) {
                uint32_t j = cand[lane];
                const float4 pj = spos[j];
                cP[lane] = make_float4(pj.x, pj.y, pj.z, 0.5f * sh[j]);
                cM[lane] = pj.w;
                cRb[lane] = pj.w > 0.f ? srb[j] : 0.0;
                note_owner(own, j, grp.x, grp.y, lane);
            }
            __syncthreads();
            // the forward's pair test (k_density), the lane's own particle excluded
            uint64_t rest = valid ? pair_mask(cP, cnt, me, qh) & ~self_bit(own, cand, cnt, lane, t) : 0ull;
            while (__ballot(rest != 0ull)) {
                if (!rest) continue;
                const int jj = __builtin_ctzll(rest);
                rest &= rest - 1ull;
                const float4 pj = cP[jj];
                const double mj = cM[jj];
                const double d[3] = {(double)me.x - (double)pj.x, (double)me.y - (double)pj.y,
                                     (double)me.z - (double)pj.z};
                const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const double irr = r2 > 0.0 ? nr_rsq(r2) : 0.0;
                const double rr = r2 * irr;
                const double inv_h = nr_rcp((double)qh + (double)pj.w);
                const double q = rr * inv_h;
                const double sig = (8.0 / 3.14159265358979323846) * (inv_h * inv_h * inv_h);
                const double w = q < 0.5 ? sig * (6.0 * (q * q * q - q * q) + 1.0)
                                         : (q < 1.0 ? sig * 2.0 * (1.0 - q) * (1.0 - q) * (1.0 - q) : 0.0);
                const double dw = q < 0.5 ? sig * inv_h * (18.0 * q * q - 12.0 * q)
                                          : (q < 1.0 ? sig * inv_h * -6.0 * (1.0 - q) * (1.0 - q) : 0.0);
                const double rbj = cRb[jj];
                const double coef = (rbi * mj + rbj * (double)me.w) * (dw * irr);
#pragma unroll
                for (int k = 0; k < 3; ++k) xb[k] += coef * d[k];
                mb += rbj * w;
            }
        });
        if (valid) {
            double4 o = sxb[t];
            o.x += xb[0];
            o.y += xb[1];
            o.z += xb[2];
            o.w += mb + (live ? rbi * (8.0 / 3.14159265358979323846) / ((double)hi * hi * hi) : 0.0);
            sxb[t] = o;
        }
    }
}

__global__ __launch_bounds__(256) void k_vjp_scatter(uint32_t n, const uint32_t* __restrict__ order,
                                                     const double4* __restrict__ sxb,
                                                     const double4* __restrict__ svb, double* xb, double* vb,
                                                     double* ub, double* mb) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t o = order[t];
    const double4 a = sxb[t], b = svb[t];
    xb[3 * o] += a.x;
    xb[3 * o + 1] += a.y;
    xb[3 * o + 2] += a.z;
    mb[o] += a.w;
    vb[3 * o] += b.x;
    vb[3 * o + 1] += b.y;
    vb[3 * o + 2] += b.z;
    ub[o] += b.w;
}

// ---- state snapshots ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_set_state(uint32_t n, const double* __restrict__ x,
                                                   const double* __restrict__ v, const float* __restrict__ u,
                                                   float4* posh, float4* velu, double* pos64, double* vel64) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = posh[i], w = velu[i];
    p.x = (float)x[3 * i];
    p.y = (float)x[3 * i + 1];
    p.z = (float)x[3 * i + 2];
    w.x = (float)v[3 * i];
    w.y = (float)v[3 * i + 1];
    w.z = (float)v[3 * i + 2];
    w.w = u[i];
    posh[i] = p;
    velu[i] = w;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pos64[3 * i + k] = x[3 * i + k];
        vel64[3 * i + k] = v[3 * i + k];
    }
}
__global__ __launch_bounds__(256) void k_save_state(uint32_t n, const double* __restrict__ pos64,
                                                    const double* __restrict__ vel64,
                                                    const float4* __restrict__ velu, double* x, double* v,
                                                    float* u) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[3 * i + k] = pos64[3 * i + k];
        v[3 * i + k] = vel64[3 * i + k];
    }
    u[i] = velu[i].w;
}

__global__ __launch_bounds__(256) void k_zero_f64(double* p, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ---- launchers -----------------------------------------------------------------------------------------
static inline uint32_t cdiv_u(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

static void zero_f64(double* p, size_t count, hipStream_t s) {
    const uint32_t blocks = (uint32_t)std::min<size_t>(2048, (count + 255) / 256);
    hipLaunchKernelGGL(k_zero_f64, dim3(std::max<uint32_t>(blocks, 1)), dim3(256), 0, s, p, count);
}

void launch_adj_loss(const Dev& d, const Adj& a, const LossParams& lp, double seed_scale, hipStream_t s) {
    const uint32_t blocks = std::min<uint32_t>(1024, cdiv_u(d.n, kRedThreads));
    zero_f64(a.red, 16, s);
    hipLaunchKernelGGL(k_loss_moments, dim3(blocks), dim3(kRedThreads), 0, s, d.n, d.pos64, d.vel64, d.mass, a.red);
    hipLaunchKernelGGL(k_loss_terms, dim3(blocks), dim3(kRedThreads), 0, s, d.n, d.pos64, d.vel64, d.mass, lp,
                       seed_scale, a.red, a.xb, a.vb, a.ub, a.mb);
    hipLaunchKernelGGL(k_loss_finish, dim3(cdiv_u(d.n, 256)), dim3(256), 0, s, d.n, d.pos64, d.vel64, d.mass, a.red,
                       seed_scale, a.xb, a.vb, a.mb);
}

void launch_vjp_kick(const Dev& d, const Adj& a, const float* u_mid, const float4* acc, float dt, hipStream_t s) {
    hipLaunchKernelGGL(k_vjp_kick, dim3(cdiv_u(d.n, 256)), dim3(256), 0, s, d.n, d.mass, u_mid, acc, dt, a.vb, a.ub,
                       a.ab);
}
void launch_vjp_kick_drift(const Dev& d, const Adj& a, const float* u_start, const float4* acc, float dt,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_vjp_kick_drift, dim3(cdiv_u(d.n, 256)), dim3(256), 0, s, d.n, d.mass, u_start, acc, dt,
                       a.xb, a.vb, a.ub, a.ab);
}

void launch_vjp_eval(const Dev& d, const StepParams& p, int gravity_mode, const Adj& a, hipStream_t s) {
    const uint32_t n = d.n;
    zero_f64(reinterpret_cast<double*>(a.sxb), 4 * (size_t)n, s);
    zero_f64(reinterpret_cast<double*>(a.svb), 4 * (size_t)n, s);
    zero_f64(a.srb, n, s);
    hipLaunchKernelGGL(k_vjp_gather, dim3(cdiv_u(n, 256)), dim3(256), 0, s, n, d.vals, d.spos, a.ab, a.sa);
    if (gravity_mode == 1) {  // direct
        hipLaunchKernelGGL(k_vjp_gravity_direct, dim3(cdiv_u(n, kDirTile)), dim3(kDirTile), 0, s, n, d.spos, a.sa,
                           p.soft2, a.sxb);
    } else if (gravity_mode == 0) {  // tree, the forward's groups (one particle per lane)
        const uint32_t nodes = 2 * d.L - 1;
        zero_f64(reinterpret_cast<double*>(a.nbar), 4 * (size_t)kNbarSlots * nodes, s);
        hipLaunchKernelGGL(k_vjp_gravity_tree, dim3(std::min<uint32_t>(kWalkWaves, n)), dim3(kWave), 0, s, d.spos, n,
                           d.K, d.L, d.nG, p.soft2, a.sa, d.groups, d.scal, a.nbar, nodes, a.sxb);
        hipLaunchKernelGGL(k_vjp_nbar_reduce, dim3(cdiv_u(nodes, 256)), dim3(256), 0, s, nodes, a.nbar);
        hipLaunchKernelGGL(k_vjp_downsweep, dim3(cdiv_u(n, 256)), dim3(256), 0, s, n, d.K, d.L, d.parent, d.nA,
                           a.nbar, d.spos, a.sxb);
    }
    hipLaunchKernelGGL(k_vjp_force, dim3(std::min<uint32_t>(kWalkWaves, n)), dim3(kWave), 0, s, d.spos, d.sh,
                       d.svel, n, d.K, d.L, d.nS, d.nrange, (uint32_t)p.sph_expand, d.srho64,
                       d.spr64, d.scs64, p.av_alpha, p.av_beta, a.sa, d.groups, d.scal, a.sxb, a.svb, a.srb,
                       a.sqb, a.scb);
    hipLaunchKernelGGL(k_vjp_eos, dim3(std::min<uint32_t>(1024, cdiv_u(n, kEosAdjThreads))), dim3(kEosAdjThreads),
                       0, s, n, d.svel, d.smat, d.eos, d.srho64, a.sqb, a.scb, a.srb, a.svb, a.pbar);
    hipLaunchKernelGGL(k_vjp_density, dim3(std::min<uint32_t>(kWalkWaves, n)), dim3(kWave), 0, s, d.spos, d.sh, n,
                       d.K, d.L, d.nS, d.nrange, (uint32_t)p.sph_expand, a.srb, d.groups, d.scal,
                       a.sxb);
    hipLaunchKernelGGL(k_vjp_scatter, dim3(cdiv_u(n, 256)), dim3(256), 0, s, n, d.vals, a.sxb, a.svb, a.xb, a.vb,
                       a.ub, a.mb);
}

void launch_set_state(const Dev& d, const double* x, const double* v, const float* u, hipStream_t s) {
    hipLaunchKernelGGL(k_set_state, dim3(cdiv_u(d.n, 256)), dim3(256), 0, s, d.n, x, v, u, d.posh, d.velu, d.pos64,
                       d.vel64);
}
void launch_save_state(const Dev& d, double* x, double* v, float* u, hipStream_t s) {
    hipLaunchKernelGGL(k_save_state, dim3(cdiv_u(d.n, 256)), dim3(256), 0, s, d.n, d.pos64, d.vel64, d.velu, x, v,
                       u);
}

}  // namespace isph
