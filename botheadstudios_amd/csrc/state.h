// Device buffers of one stepper handle and the launcher interface between the host code and the kernels.
#pragma once

#include <algorithm>
#include <string>

#include "common.h"

namespace isph {

// Upsweep chunking: level-1 chunks of kUpLeaves leaves; each chunk carries at most kUpCarry ids of
// higher-level nodes to the next level.
constexpr uint32_t kUpLeaves = 512;
#ifndef ISPH_SORT_TILE
#define ISPH_SORT_TILE 4096
#endif
constexpr int kSortTile = ISPH_SORT_TILE;  // radix sort: keys per workgroup per pass (256 threads)
constexpr uint32_t kUpCarry = 256;

// SPH walk record of a node, 48 B (one cache line for most nodes): the box with the largest smoothing
// length, then for an internal node its children and leaf range.
struct SphNode {
    float4 lo;  // bmin.xyz, max h
    float4 hi;  // bmax.xyz, MAC radius size/theta + |centre - com|
    uint4 lr;   // left child, right child, first leaf, last leaf
};

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
    double* acc64 = nullptr;  // mixed precision: ax, ay, az of the last evaluation, unrounded (3n)

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
    uint32_t* cslot = nullptr;  // upsweep: LDS slot of each internal node within its level's chunk
    uint2* nrange = nullptr;    // [first, last] leaf range of each internal node
    uint8_t* nlevel = nullptr;  // upsweep level of each internal node
    uint32_t* carry = nullptr;   // upsweep: higher-level node ids carried per chunk (two alternating lists)
    uint32_t* ncarry = nullptr;  // upsweep: their counts
    uint32_t carry_blocks = 0;   // chunks per list (level 1: kUpLeaves leaves each)
    uint2* groups = nullptr;    // SPH walk groups (first sorted particle, count), Morton order
    uint2* ggroups = nullptr;   // gravity walk groups (up to 64 x particles-per-lane), Morton order
    // Candidate lists of the SPH walk groups, written by the density walk and read back by the force pass
    // (same groups, same candidates, same order): per group the first chunk, per chunk its count, the next
    // chunk and 64 candidate slots. Chunks come in blocks of kChunksPerBlock taken by one atomic.
    uint32_t* chead = nullptr;
    uint32_t* cnext = nullptr;
    uint32_t* ccnt = nullptr;
    uint32_t* cpool = nullptr;
    uint64_t* cmask = nullptr;  // per chunk and group lane: the density's pair mask over the chunk's slots
    uint32_t cblocks = 0;
    uint32_t* gflag = nullptr;  // group marks at their first particle (all zero between builds)
    uint32_t* gscan = nullptr;  // per-64-particle mark counts / offsets
    uint4* nG = nullptr;   // gravity walk record, 2 x 16 B: (com.xyz, mass), (MAC radius bits, left, right, 0)
    SphNode* nS = nullptr;  // SPH walk record: box, largest h, children and leaf range (one load per walk step)
    float4* nA = nullptr;   // com.xyz, mass
    float4* nQ0 = nullptr; // quadrupole xx, yy, zz, xy
    float2* nQ1 = nullptr; // quadrupole xz, yz

    // f64 reference form: residuals of the f64 state against the f32 copies the walks read (x64 = x32 + lo),
    // sorted; node depths, the f64 COM and mass of every internal node, the f64 gravity of each sorted particle
    float4* slo = nullptr;      // position residual xyz (w unused)
    float4* svlo = nullptr;     // velocity residual xyz (w unused)
    uint8_t* ndepth = nullptr;  // [L-1]
    double4* nA64 = nullptr;    // [L-1] com.xyz, mass
    double4* sgrav64 = nullptr; // xyz (w unused)

    float4* ext = nullptr;       // [kMaxExtBodies] external point masses (xyz, mass)
    uint32_t* mtarget = nullptr; // merge target of each particle (upload order)

    // ensembles (ninst > 1): instance k owns particles [ioff[k], ioff[k+1]) in upload and sorted order
    uint32_t ninst = 1;
    uint32_t* inst = nullptr;   // instance of each particle, upload order
    uint32_t* mcode = nullptr;  // Morton code against the instance box, upload order
    uint64_t* lkey = nullptr;   // leaf keys (instance << 30 | code), sorted order
    uint32_t* ioff = nullptr;   // [ninst + 1]
    uint32_t* ibox = nullptr;   // [6 ninst] bbox keys
    uint32_t* iroot = nullptr;  // [ninst] root node of each instance's subtree
    uint32_t* isig = nullptr;   // [ninst] Courant signal minimum (float bits)
    float* idt = nullptr;       // [ninst] time step
    double* itime = nullptr;    // [ninst] simulated time
    // sorted instance ids (the sort keys after the instance pass), or nullptr for one system
    const uint32_t* sinst() const { return ninst > 1 ? keys : nullptr; }

    // scalars
    // [0..5] bbox keys, [6] Courant signal min (float bits), [7] crossing-node count, [8] SPH walk groups,
    // [12] gravity walk groups, [512..) / [1024..) / [1536..) work-queue areas of the gravity, density and
    // force walks
    uint32_t* scal = nullptr;
    float* dt = nullptr;       // current dt
    double* time = nullptr;    // accumulated simulated time
    unsigned long long* counters = nullptr;  // [0] cells, [1] particles, [2] quad cells, [3] dens pairs, [4] force pairs
    double* red = nullptr;     // reduction partials (diagnostics)
    double* inject = nullptr;  // energy added by the u >= 0 clamp of the kicks since the last reset (J)

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
    float group_extent = 4.0f;  // walk groups: box diagonal <= this many times the largest h
    int grav_ppl = 1;           // gravity walk: particles per lane (1 or 2)
    int sph_expand = 4;         // SPH walks take subtrees of <= this many particles whole as candidates
    // un-resolved bulk (sph_step.wgsl:120-162): bulk_cr = (centre, R_core), bulk_vm = (velocity, mass);
    // R_core == 0 none, < 0 planar (bulk_cr.xyz = plane point, bulk_vm = (up, g))
    float4 bulk_cr = {0.f, 0.f, 0.f, 0.f};
    float4 bulk_vm = {0.f, 0.f, 0.f, 0.f};
    uint32_t n_ext = 0;          // external point masses (sph_step.wgsl:249-254)
    float omega = 0.0f;          // relaxation frame spin about +z (sph_step.wgsl:353)
    uint32_t merge_budget = 0;   // pairwise merging while n > budget (sph_step.wgsl:415-541); 0 = never
    // candidate-cache blocks the density walk may use (capped at what is allocated); a group whose list
    // does not fit is walked again by the force pass, so a low cap exercises that path
    uint32_t cache_blocks = 0xffffffffu;
    int ref64 = 0;  // the f64 reference form (mixed as well): f64 coordinates in every pair term, f64 gravity
    double soft2_64 = 0.0;  // softening^2 in f64 (the reference form's gravity)
};

constexpr uint32_t kMaxExtBodies = 32;  // gpu_sph.rs:155
constexpr uint32_t kNoMerge = 0xffffffffu;

// What a force evaluation adds beyond self-gravity and SPH: the bulk and the external masses.
struct Boundary {
    float4 bulk_cr, bulk_vm;
    const float4* ext;
    uint32_t n_ext;
    float soft2;
};

constexpr uint32_t kGroupMax = 64;    // particles per SPH walk group (one per lane)
constexpr uint32_t kWalkWaves = 8192; // resident waves of the work-queue walks
// Statistics counters: kCounterSlots per walk wave (one-wave workgroup w owns counters[w * kCounterSlots ..]),
// so the walks update them without atomics; the host sums them when asked.
constexpr uint32_t kCounterSlots = 16;
#ifndef ISPH_QUEUES
#define ISPH_QUEUES 8
#endif
#ifndef ISPH_QUEUE_STRIDE
#define ISPH_QUEUE_STRIDE 32
#endif
#ifndef ISPH_QUEUE_MASK
#define ISPH_QUEUE_MASK 1
#endif
constexpr uint32_t kQueues = ISPH_QUEUES;             // work queues per walk (1 or 8)
constexpr uint32_t kQueueStride = ISPH_QUEUE_STRIDE;  // uint32 slots between queue heads (32 = 128 B line)
static_assert(kQueues == 1 || kQueues == 8, "one queue or one per XCD");
static_assert(kQueueStride >= 1 && kQueueStride * 8 < 512, "queue area overflow");
constexpr int kQueueArea = 512;  // per walk: kQueues heads, then the exhausted-range mask
constexpr int kGroupsSph = 8, kGroupsGrav = 12;
constexpr int kPoolHead = 20;                 // scalar slot: next free candidate-chunk block
constexpr uint32_t kChunksPerBlock = 32;
constexpr uint32_t kNoChunk = 0xffffffffu;
constexpr uint32_t kChunkOverflow = 0xfffffffeu;  // the group's list did not fit: the force pass walks
constexpr int kQueueGrav = 512, kQueueDensity = 1024, kQueueForce = 1536;
constexpr int kScalars = 2048;

// Work queues over a Morton-ordered group list. With one queue per XCD the list is cut into 8 contiguous
// ranges; a wave starts on the range of the XCD it was dispatched to (workgroup id mod 8) and moves on to
// the next range when its own runs dry, so concurrently walked groups on one XCD are spatial neighbours and
// share its L2. The wave that drains a range sets its bit in the area's mask, and later waves skip that
// range with one load instead of a failing atomic. next() returns false once every range is drained, an
// exit every wave reaches.
// The first `statics` groups of a wave's own range are assigned without atomics (wave j of the W waves
// dispatched to that range takes lo + j, lo + j + W, ...); the range's counter hands out the rest.
struct WalkQueue {
    uint32_t* area;
    uint32_t total;
    uint32_t q, tries, taken, statics;
    __device__ WalkQueue(uint32_t* a, uint32_t total_groups, uint32_t static_takes = 0)
        : area(a), total(total_groups), statics(static_takes) {
        q = kQueues == 1 ? 0u : (blockIdx.x & 7u);
        tries = 0;
        taken = 0;
    }
    __device__ __forceinline__ bool next(uint32_t& gi) {
        uint32_t* mask = area + kQueueStride * kQueues;
        while (tries < kQueues) {
            const uint32_t lo = (uint32_t)(((uint64_t)q * total) / kQueues);
            const uint32_t hi = (uint32_t)(((uint64_t)(q + 1) * total) / kQueues);
            const uint32_t waves = kQueues == 1 ? gridDim.x : (gridDim.x + 7u - q) / 8u;  // this range's waves
            const uint32_t first = lo + statics * waves;  // first group the counter hands out
            if (taken < statics) {
                const uint32_t g = lo + (kQueues == 1 ? blockIdx.x : blockIdx.x / 8u) + taken * waves;
                ++taken;
                if (g < hi) {
                    gi = g;
                    return true;
                }
                taken = statics;
            }
            uint32_t g = hi;
            if (__lane_id() == 0) {
                bool drained = false;
                if (ISPH_QUEUE_MASK && kQueues > 1)
                    drained = (__hip_atomic_load(mask, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> q) & 1u;
                if (!drained) {
                    g = first + atomicAdd(&area[q * kQueueStride], 1u);
                    if (ISPH_QUEUE_MASK && kQueues > 1 && g >= hi) atomicOr(mask, 1u << q);
                }
            }
            g = __builtin_amdgcn_readfirstlane(g);
            if (g < hi) {
                gi = g;
                return true;
            }
            q = (q + 1) % kQueues;
            ++tries;
        }
        return false;
    }
};

// ---- launchers (each .hip file implements its own) ----
void launch_reset_scalars(const Dev& d, hipStream_t s, int mode = 0);
void launch_bbox(const Dev& d, const float4* pos, hipStream_t s);
void launch_morton(const Dev& d, const float4* pos, hipStream_t s);
void launch_radix_sort(const Dev& d, hipStream_t s);  // 30-bit keys, 4 passes; result in keys/vals
void launch_gather(const Dev& d, const float4* pos, const float* mass, hipStream_t s);
void launch_gather_vel(const Dev& d, hipStream_t s);

void launch_radix_pass(const Dev& d, const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                       int shift, hipStream_t s);

// ensembles
void launch_reset_seg(const Dev& d, hipStream_t s, int mode);
void launch_sorted_seg(const Dev& d, const float4* pos, const float* mass, hipStream_t s);
void launch_inst_roots(const Dev& d, hipStream_t s);
void launch_set_dt_cfl_seg(const Dev& d, float cfl, hipStream_t s);
void launch_set_dt_value_seg(const Dev& d, float dt, hipStream_t s);
void launch_advance_time_seg(const Dev& d, hipStream_t s);

void launch_karras(const Dev& d, hipStream_t s);
void launch_upsweep(const Dev& d, const StepParams& p, hipStream_t s);
void launch_groups(const Dev& d, const StepParams& p, uint32_t maxc, uint2* out, uint32_t* total, hipStream_t s);
void launch_exclusive_scan(uint32_t* data, uint32_t count, hipStream_t s);
// One stable 8-bit LSD pass over n (key, value) pairs; hist holds 256 * ceil(n / 4096) counters.
void launch_radix_pass_n(uint32_t n, uint32_t* hist, const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                         uint32_t* vout, int shift, hipStream_t s);

void launch_gravity_tree(const Dev& d, const StepParams& p, hipStream_t s);
void launch_gravity_direct(const Dev& d, const StepParams& p, hipStream_t s);
// f64 reference form: residuals of the sorted state, node COM and mass in f64 (level by level from the
// deepest), and the f64 gravity (tree walk with the f32 walk's acceptance decisions, or the exact sum)
void launch_gather_lo(const Dev& d, hipStream_t s);
void launch_com64(const Dev& d, hipStream_t s);
void launch_gravity_tree_ref(const Dev& d, const StepParams& p, hipStream_t s);
void launch_gravity_direct_ref(const Dev& d, const StepParams& p, hipStream_t s);
void launch_gravity_zero_ref(const Dev& d, hipStream_t s);
// The exact sum on explicit arrays: bodies (x, y, z, m) in, kG-scaled accelerations out, compensated sums
// when `comp`; no handle, no allocation.
void launch_gravity_direct_raw(const float4* bodies, uint32_t n, float soft2, float4* out, bool comp, hipStream_t s);
// The calling thread's error message for entry points that take no handle (read by isph_last_error(NULL)).
void set_thread_error(const std::string& msg);
void launch_gravity_zero(const Dev& d, hipStream_t s);
void launch_scatter_gravity(const Dev& d, hipStream_t s);  // sgrav (sorted) -> acc (orig), for the gravity API

void launch_density(const Dev& d, const StepParams& p, hipStream_t s);
void launch_eos(const Dev& d, hipStream_t s);
void launch_force(const Dev& d, const StepParams& p, hipStream_t s);

void launch_kick_drift(const Dev& d, const StepParams& p, hipStream_t s);
void launch_relax(const Dev& d, const StepParams& p, float damp, hipStream_t s);
void launch_merge(const Dev& d, const StepParams& p, hipStream_t s);  // pick, apply, retire (after density)
void launch_kick(const Dev& d, hipStream_t s);
void launch_set_dt_cfl(const Dev& d, float cfl, hipStream_t s);
void launch_set_dt_value(const Dev& d, float dt, hipStream_t s);
void launch_advance_time(const Dev& d, hipStream_t s);

void launch_unpack(const Dev& d, uint32_t n, hipStream_t s);
void launch_pack(const Dev& d, hipStream_t s);
void launch_sync64_from32(const Dev& d, hipStream_t s);

void launch_diag(const Dev& d, float soft, int with_pe, hipStream_t s);  // writes d.red[0..7]

}  // namespace isph
