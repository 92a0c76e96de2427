// Shared device definitions for the gfx950 particle-step kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace isph {

constexpr int kMaxMaterials = 16;           // gpu_sph.rs:156 MAX_MATERIALS
constexpr uint32_t kNoParent = 0xffffffffu; // bh_gravity.wgsl:32
constexpr float kG = 6.674e-11f;            // sph_step.wgsl:21, bh_gravity.wgsl:11
constexpr double kG64 = 6.674e-11;          // orbit.rs:18
constexpr float kPi = 3.14159265359f;       // sph_step.wgsl:20
constexpr int kWave = 64;

// Algorithmic flop constants per counted interaction (the roofline numerator).
constexpr double kFlopsMonopole = 20.0;  // d(3) r2(6) rsqrt(1) r^-3(3) m*r^-3(1) accumulate(6)
constexpr double kFlopsQuadrupole = 51.0;// d(3) r2(6) rsqrt(1) inv2(1) w(2) Q.d(15) tq(7) s(2) terms(8) acc(6)
constexpr double kFlopsDensityPair = 20.0;
constexpr double kFlopsForcePair = 75.0;

// Tillotson parameters (eos.rs:27-47 field order), f64.
struct Till {
    double rho0, a, b, cap_a, cap_b, e0, e_iv, e_cv, alpha, beta;
};

struct EosTable {
    Till m[kMaxMaterials];
};

// ---- wave64 helpers --------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}
__device__ __forceinline__ int prefix_count(uint64_t mask) { return __popcll(mask & lanemask_lt()); }

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
// A value every lane holds alike, moved to a scalar register.
__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Order-preserving float <-> uint map of bh_gravity.wgsl:49-57.
__device__ __forceinline__ uint32_t bbox_key(float f) {
    uint32_t k = __float_as_uint(f);
    return (k >> 31) ? ~k : (k ^ 0x80000000u);
}
__device__ __forceinline__ float bbox_decode(uint32_t k) {
    return __uint_as_float((k >> 31) ? (k ^ 0x80000000u) : ~k);
}

// ---- forward-mode dual numbers (value + N partials), for the adjoint's per-particle EOS Jacobian --------
template <int N>
struct Dual {
    double v;
    double d[N];
};
template <int N>
__device__ __forceinline__ Dual<N> dual_const(double x) {
    Dual<N> r;
    r.v = x;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dual_var(double x, int k) {
    Dual<N> r = dual_const<N>(x);
    r.d[k] = 1.0;
    return r;
}
#define ISPH_DUAL_BINOP(OP, VEXPR, DEXPR)                                                                 \
    template <int N>                                                                                      \
    __device__ __forceinline__ Dual<N> operator OP(const Dual<N>& a, const Dual<N>& b) {                  \
        Dual<N> r;                                                                                        \
        r.v = VEXPR;                                                                                      \
        _Pragma("unroll") for (int k = 0; k < N; ++k) r.d[k] = DEXPR;                                     \
        return r;                                                                                         \
    }                                                                                                     \
    template <int N>                                                                                      \
    __device__ __forceinline__ Dual<N> operator OP(const Dual<N>& a, double bv) {                         \
        return a OP dual_const<N>(bv);                                                                    \
    }                                                                                                     \
    template <int N>                                                                                      \
    __device__ __forceinline__ Dual<N> operator OP(double av, const Dual<N>& b) {                         \
        return dual_const<N>(av) OP b;                                                                    \
    }
ISPH_DUAL_BINOP(+, a.v + b.v, a.d[k] + b.d[k])
ISPH_DUAL_BINOP(-, a.v - b.v, a.d[k] - b.d[k])
ISPH_DUAL_BINOP(*, a.v * b.v, a.d[k] * b.v + a.v * b.d[k])
ISPH_DUAL_BINOP(/, a.v / b.v, (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v))
#undef ISPH_DUAL_BINOP
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
    return dual_const<N>(0.0) - a;
}
template <int N>
__device__ __forceinline__ bool operator>=(const Dual<N>& a, double b) { return a.v >= b; }
template <int N>
__device__ __forceinline__ bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N>
__device__ __forceinline__ bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }

// Scalar functions spelled once for double and for duals (branch decisions follow the value).
__device__ __forceinline__ double gexp(double x) { return exp(x); }
__device__ __forceinline__ double gsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double gfabs(double x) { return fabs(x); }
__device__ __forceinline__ double gfmax(double x, double y) { return fmax(x, y); }
__device__ __forceinline__ double gval(double x) { return x; }
template <int N>
__device__ __forceinline__ Dual<N> gexp(const Dual<N>& a) {
    Dual<N> r;
    r.v = exp(a.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = r.v * a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> gsqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = sqrt(a.v);
    const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> gfabs(const Dual<N>& a) {
    return a.v < 0.0 ? -a : a;
}
template <int N>
__device__ __forceinline__ Dual<N> gfmax(const Dual<N>& a, double y) {
    return a.v >= y ? a : dual_const<N>(y);
}
template <int N>
__device__ __forceinline__ double gval(const Dual<N>& a) { return a.v; }

// ---- Tillotson (eos.rs:52-96; the GPU shader's f32 form is sph_step.wgsl:87-109) ----------------------
// Written once for a scalar type T: double for the step, Dual<N> for the adjoint's Jacobian.
template <class T>
struct TillT {
    T rho0, a, b, cap_a, cap_b, e0, e_iv, e_cv, alpha, beta;
};
template <class T>
__device__ __forceinline__ T till_pressure_t(const TillT<T>& t, T rho_in, T u) {
    T rho = gfmax(rho_in, 1.0e-9);
    T eta = rho / t.rho0;
    T mu = eta - 1.0;
    T omega = u / (t.e0 * eta * eta) + 1.0;
    T pc = (t.a + t.b / omega) * rho * u + t.cap_a * mu + t.cap_b * mu * mu;
    if (eta >= 1.0 || u <= t.e_iv) return pc;
    T z = t.rho0 / rho - 1.0;
    T pe = t.a * rho * u + (t.b * rho * u / omega + t.cap_a * mu * gexp(-t.beta * z)) * gexp(-t.alpha * z * z);
    if (u >= t.e_cv) return pe;
    return ((u - t.e_iv) * pe + (t.e_cv - u) * pc) / (t.e_cv - t.e_iv);
}
// Pressure and squared sound speed (central differences, relative step 1e-4 as eos.rs:88-96).
template <class T>
__device__ __forceinline__ void till_p_c2_t(const TillT<T>& t, T rho_in, T u, T& p, T& c2) {
    T rho = gfmax(rho_in, 1.0e-9);
    T dr = rho * 1.0e-4;
    T du = gfabs(u) * 1.0e-4 + 1.0;
    T dp_drho = (till_pressure_t(t, rho + dr, u) - till_pressure_t(t, rho - dr, u)) / (2.0 * dr);
    T dp_du = (till_pressure_t(t, rho, u + du) - till_pressure_t(t, rho, u - du)) / (2.0 * du);
    p = till_pressure_t(t, rho, u);
    c2 = gfmax(dp_drho + p / (rho * rho) * dp_du, 0.0);
}
__device__ __forceinline__ double till_pressure(const Till& t, double rho_in, double u) {
    return till_pressure_t<double>(reinterpret_cast<const TillT<double>&>(t), rho_in, u);
}
__device__ __forceinline__ void till_p_c2(const Till& t, double rho_in, double u, double& p, double& c2) {
    till_p_c2_t<double>(reinterpret_cast<const TillT<double>&>(t), rho_in, u, p, c2);
}

// ---- cubic spline with support h (sph_step.wgsl:73-86, atmosphere.rs:52-75) ---------------------------
// W(r,h) and dW/dr from q = r/h and 1/h.
// Hardware square root and reciprocal (1 ulp) for the plain form's pair terms: the correctly rounded
// library forms expand into refinement sequences several times longer.
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// f64 reciprocal and reciprocal square root: the hardware estimate refined by one Newton-Raphson step,
// which squares its relative error. Measured on gfx950 over 4 Mi inputs in [2^-100, 2^100]: estimates
// within 4.6e-8 (rcp) and 5.2e-8 (rsq), refined within 2.2e-15 and 4.3e-15 of the correctly rounded
// results. The mixed form's pair terms need f64 range and cancellation-free sums, not the last ulp: the
// f32-stored coordinates limit its per-particle error near 1e-7. A compiled f64 division or sqrt costs
// about three times as many instructions.
__device__ __forceinline__ double nr_rcp(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    return fma(y, fma(-x, y, 1.0), y);
}
__device__ __forceinline__ double nr_rsq(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    return fma(y, fma(-0.5 * x * y, y, 0.5), y);
}

__device__ __forceinline__ float spline_w(float q, float inv_h) {
    float sig = (8.0f / kPi) * inv_h * inv_h * inv_h;
    if (q < 0.5f) return sig * (1.0f - 6.0f * q * q + 6.0f * q * q * q);
    if (q < 1.0f) {
        float t = 1.0f - q;
        return sig * 2.0f * t * t * t;
    }
    return 0.0f;
}
__device__ __forceinline__ float spline_dw(float q, float inv_h) {
    float sig = (8.0f / kPi) * inv_h * inv_h * inv_h * inv_h;
    if (q < 0.5f) return sig * (-12.0f * q + 18.0f * q * q);
    if (q < 1.0f) {
        float t = 1.0f - q;
        return sig * (-6.0f * t * t);
    }
    return 0.0f;
}

// ---- the un-resolved bulk and its floor (sph_step.wgsl:120-162) ----------------------------------------
// Gravity of the bulk: monopole outside R_core, linear inside (-> 0 at the centre); planar mode: uniform g
// along -up. f32 as the shader.
__device__ __forceinline__ float3 bulk_gravity(float3 pos, float4 cr, float4 vm) {
    const float R = cr.w;
    if (R == 0.0f) return make_float3(0.f, 0.f, 0.f);
    if (R < 0.0f) {
        const float il = rsqrtf(vm.x * vm.x + vm.y * vm.y + vm.z * vm.z);
        return make_float3(-vm.x * il * vm.w, -vm.y * il * vm.w, -vm.z * il * vm.w);
    }
    const float dx = pos.x - cr.x, dy = pos.y - cr.y, dz = pos.z - cr.z;
    const float r = sqrtf(dx * dx + dy * dy + dz * dz);
    if (r < 1.0f) return make_float3(0.f, 0.f, 0.f);
    const float gg = r >= R ? kG * vm.w / (r * r) : kG * vm.w * r / (R * R * R);
    return make_float3(-dx / r * gg, -dy / r * gg, -dz / r * gg);
}
// The non-injecting floor: a particle below the plane / inside R_core is projected onto it and its
// velocity into the floor (relative to the bulk) removed. T is float or double.
template <class T>
__device__ __forceinline__ void bulk_floor(T& x, T& y, T& z, T& vx, T& vy, T& vz, float4 cr, float4 vm) {
    const float R = cr.w;
    if (R == 0.0f) return;
    if (R < 0.0f) {
        const T il = (T)1 / sqrt((T)vm.x * vm.x + (T)vm.y * vm.y + (T)vm.z * vm.z);
        const T ux = vm.x * il, uy = vm.y * il, uz = vm.z * il;
        const T depth = (x - cr.x) * ux + (y - cr.y) * uy + (z - cr.z) * uz;
        if (depth < (T)0) {
            x -= ux * depth;
            y -= uy * depth;
            z -= uz * depth;
            const T vn = vx * ux + vy * uy + vz * uz;
            if (vn < (T)0) {
                vx -= ux * vn;
                vy -= uy * vn;
                vz -= uz * vn;
            }
        }
        return;
    }
    const T dx = x - cr.x, dy = y - cr.y, dz = z - cr.z;
    const T r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r >= (T)R || r < (T)1) return;
    const T nx = dx / r, ny = dy / r, nz = dz / r;
    x = cr.x + nx * R;
    y = cr.y + ny * R;
    z = cr.z + nz * R;
    const T vn = (vx - vm.x) * nx + (vy - vm.y) * ny + (vz - vm.z) * nz;
    if (vn < (T)0) {
        vx -= nx * vn;
        vy -= ny * vn;
        vz -= nz * vn;
    }
}

// Compensated accumulation (Neumaier/Kahan) for the mixed-precision form.
struct KahanF {
    float s, c;
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// Map a linear block id so that consecutive work lands on the same XCD (8 XCDs, round-robin dispatch):
// the blocks dispatched to XCD x receive the x-th contiguous slice of the work.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    uint32_t per = nblocks / kXcd;
    uint32_t full = per * kXcd;
    if (bid >= full) return bid;
    return (bid % kXcd) * per + bid / kXcd;
}

}  // namespace isph
