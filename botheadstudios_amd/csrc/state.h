// Device buffers of one stepper handle and the launcher interface between the host code and the kernels.
#pragma once

#include <algorithm>

#include "common.h"

namespace isph {

// All device arrays. "orig" arrays are in upload order; "sorted" arrays in Morton order of the last
// evaluation (order[t] = upload index of sorted slot t).
struct Dev {
    uint32_t n = 0;     // live particles
    uint32_t cap = 0;   // capacity
    uint32_t K = 1;     // leaf bucket size
    uint32_t L = 1;     // leaves = ceil(n / K)

    // master state, upload order
    float4* posh = nullptr;   // x, y, z, h
    float4* velu = nullptr;   // vx, vy, vz, u
    float* mass = nullptr;
    uint32_t* mat = nullptr;
    uint32_t* prov = nullptr;
    float* rho = nullptr;
    float4* acc = nullptr;    // ax, ay, az, du/dt of the last evaluation
    double* pos64 = nullptr;  // mixed precision: x, y, z per particle (3n)
    double* vel64 = nullptr;  // mixed precision: vx, vy, vz per particle (3n)

    // Morton + sort
    uint32_t* keys = nullptr;  // codes
    uint32_t* vals = nullptr;  // order
    uint32_t* keys_alt = nullptr;
    uint32_t* vals_alt = nullptr;
    uint32_t* hist = nullptr;  // 256 * sort blocks
    uint32_t nsort_blocks = 0;

    // sorted working copies
    float4* spos = nullptr;  // x, y, z, mass
    float4* svel = nullptr;  // vx, vy, vz, u
    float* sh = nullptr;
    uint32_t* smat = nullptr;
    float* srho = nullptr;
    float* spr = nullptr;    // P / rho^2
    float* scs = nullptr;    // sound speed
    double* srho64 = nullptr;  // mixed precision: f64 density and P / rho^2 (sorted)
    double* spr64 = nullptr;
    double* scs64 = nullptr;
    float4* sgrav = nullptr; // gravitational acceleration, sorted
    uint32_t* lcode = nullptr;  // leaf cluster codes [L]

    // tree: internal [0, L-1), leaves [L-1, 2L-1)
    uint32_t* left = nullptr;
    uint32_t* right = nullptr;
    uint32_t* parent = nullptr;
    uint32_t* ready = nullptr;
    uint2* nrange = nullptr;    // [first, last] leaf range of each internal node
    uint32_t* cross = nullptr;  // internal nodes whose range crosses an upsweep chunk boundary
    uint2* groups = nullptr;    // walk groups: (first sorted particle, count); chunk slots then seam pieces
    uint32_t* gflag = nullptr;  // 64-particle chunks that straddle a Morton seam
    uint4* nG = nullptr;   // gravity walk record, 2 x 16 B: (com.xyz, mass), (MAC radius bits, left, right, 0)
    uint2* nLR = nullptr;  // children of internal nodes, for the SPH walks
    float4* nA = nullptr;  // com.xyz, mass
    float4* nB = nullptr;  // bmin.xyz, max h
    float4* nC = nullptr;  // bmax.xyz, MAC radius size/theta + |centre - com|
    float4* nQ0 = nullptr; // quadrupole xx, yy, zz, xy
    float2* nQ1 = nullptr; // quadrupole xz, yz

    // scalars
    // [0..5] bbox keys, [6] Courant signal min (float bits), [7] crossing-node count, [8] walk groups,
    // [9..11] work-queue heads of the gravity, density and force walks
    uint32_t* scal = nullptr;
    float* dt = nullptr;       // current dt
    double* time = nullptr;    // accumulated simulated time
    unsigned long long* counters = nullptr;  // [0] cells, [1] particles, [2] quad cells, [3] dens pairs, [4] force pairs
    double* red = nullptr;     // reduction partials (diagnostics)

    Till* eos = nullptr;       // EOS table [kMaxMaterials]

    // read-back / upload staging
    void* aos = nullptr;       // device AoS records
};

struct StepParams {
    float soft2 = 0.0f;
    float theta = 0.5f;
    float av_alpha = 1.0f;
    float av_beta = 2.0f;
    int quad = 0;
    int mixed = 0;
    float group_extent = 4.0f;  // subtree walk groups: box diagonal <= this many times the largest h
    float chunk_extent = 0.0f;  // 64-particle chunk groups: box diagonal limit in the same unit (0 = off)
};

constexpr uint32_t kGroupMax = 64;  // particles per walk group (one per lane)
constexpr uint32_t kWalkWaves = 8192; // resident waves of the work-queue walks

// ---- launchers (each .hip file implements its own) ----
void launch_reset_scalars(const Dev& d, hipStream_t s, int mode = 0);
void launch_bbox(const Dev& d, const float4* pos, hipStream_t s);
void launch_morton(const Dev& d, const float4* pos, hipStream_t s);
void launch_radix_sort(const Dev& d, hipStream_t s);  // 30-bit keys, 4 passes; result in keys/vals
void launch_gather(const Dev& d, const float4* pos, const float* mass, hipStream_t s);
void launch_gather_vel(const Dev& d, hipStream_t s);

void launch_karras(const Dev& d, hipStream_t s);
void launch_upsweep(const Dev& d, const StepParams& p, hipStream_t s);
void launch_groups(const Dev& d, const StepParams& p, hipStream_t s);
uint32_t walk_chunk_slots(const Dev& d, const StepParams& p);  // leading chunk slots of the group list

void launch_gravity_tree(const Dev& d, const StepParams& p, hipStream_t s);
void launch_gravity_direct(const Dev& d, const StepParams& p, hipStream_t s);
void launch_gravity_zero(const Dev& d, hipStream_t s);
void launch_scatter_gravity(const Dev& d, hipStream_t s);  // sgrav (sorted) -> acc (orig), for the gravity API

void launch_density(const Dev& d, const StepParams& p, hipStream_t s);
void launch_eos(const Dev& d, hipStream_t s);
void launch_force(const Dev& d, const StepParams& p, hipStream_t s);

void launch_kick_drift(const Dev& d, hipStream_t s);
void launch_kick(const Dev& d, hipStream_t s);
void launch_set_dt_cfl(const Dev& d, float cfl, hipStream_t s);
void launch_set_dt_value(const Dev& d, float dt, hipStream_t s);
void launch_advance_time(const Dev& d, hipStream_t s);

void launch_unpack(const Dev& d, uint32_t n, hipStream_t s);
void launch_pack(const Dev& d, hipStream_t s);
void launch_sync64_from32(const Dev& d, hipStream_t s);

void launch_diag(const Dev& d, float soft, int with_pe, hipStream_t s);  // writes d.red[0..7]

}  // namespace isph
