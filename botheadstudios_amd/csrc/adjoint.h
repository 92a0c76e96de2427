// Adjoint (vector-Jacobian product) of the KDK step: device buffers and launchers.
//
// The forward step is the mixed form of the stepper (f64 position / velocity state, f64 SPH pair sums,
// f32 gravity pairs). The adjoint runs it backwards one evaluation at a time: for a force evaluation
// with outputs (a_i, du_i/dt) and their adjoints (abar_i, ebar_i) it accumulates the adjoints of the
// inputs (positions, velocities, internal energies, masses) and of the Tillotson parameters. The tree
// topology, the walk groups and every acceptance decision of the gravity walk are those of the forward
// evaluation (frozen per evaluation); the Courant time steps are those of the forward run.
#pragma once

#include "state.h"

namespace isph {

// Adjoint arrays. "Sorted" arrays follow the Morton order of the evaluation being differentiated.
struct Adj {
    // upload order (the adjoint state carried between steps)
    double* xb = nullptr;   // 3n: d loss / d position
    double* vb = nullptr;   // 3n: d loss / d velocity
    double* ub = nullptr;   // n:  d loss / d u
    double* mb = nullptr;   // n:  d loss / d mass
    double4* ab = nullptr;  // n:  adjoint of one evaluation's outputs (ax, ay, az, du/dt)
    // sorted order, per evaluation
    double4* sa = nullptr;   // gathered (abar, ebar)
    double4* sxb = nullptr;  // (xbar, mbar)
    double4* svb = nullptr;  // (vbar, ubar)
    double* srb = nullptr;   // rho bar
    double* sqb = nullptr;   // (P / rho^2) bar
    double* scb = nullptr;   // sound-speed bar
    double4* nbar = nullptr; // per tree node: (COM bar, mass bar), kNbarSlots copies (slot 0 holds the sum)
    double* pbar = nullptr;  // kMaxMaterials x 4: Tillotson a, b, A, B bars
    double* red = nullptr;   // loss reductions (16 doubles)
};

// The gravity adjoint's node sums arrive as f64 atomics from every group that accepted the node; the top
// nodes take thousands each. Waves add into one of kNbarSlots copies (by workgroup id), which a reduction
// folds into slot 0 before the push-down, so each address sees an eighth of the contention.
constexpr uint32_t kNbarSlots = 8;

// Loss: sum_i m_i sig(k_r (r_i / R - r_cut)) sig(k_j (j_i / sqrt(G M r_i) - 1)), with r_i, j_i the
// distance and the specific angular momentum along `axis` about the centre of mass (total M), in f64.
struct LossParams {
    double radius, r_cut, k_r, k_j;
    double ax, ay, az;
};

void launch_adj_loss(const Dev& d, const Adj& a, const LossParams& lp, double seed_scale, hipStream_t s);
// Kick adjoints (upload order); write a.ab for the evaluation that produced `acc`.
void launch_vjp_kick(const Dev& d, const Adj& a, const float* u_mid, const float4* acc, float dt, hipStream_t s);
void launch_vjp_kick_drift(const Dev& d, const Adj& a, const float* u_start, const float4* acc, float dt,
                           hipStream_t s);
// The adjoint of one force evaluation whose sorted intermediates are in `d` (mixed form).
void launch_vjp_eval(const Dev& d, const StepParams& p, int gravity_mode, const Adj& a, hipStream_t s);
// Restore / save a state snapshot: pos64, vel64 (3n each), u (n) <-> the stepper's state arrays.
void launch_set_state(const Dev& d, const double* x, const double* v, const float* u, hipStream_t s);
void launch_save_state(const Dev& d, double* x, double* v, float* u, hipStream_t s);

}  // namespace isph
