/*
 * isph.h — C-ABI of the MI355X particle physics step (LBVH Barnes-Hut gravity + SPH/Tillotson + KDK).
 *
 * Drop-in boundary for integrity-engine's GPU SPH stepper and GPU gravity. Each entry point names the
 * reference interface it replaces (paths relative to integrity-engine/). The wgpu `&Device`, `&Queue` and
 * `&mut CommandEncoder` arguments of the reference cannot cross to HIP: they are replaced by the opaque
 * handle, which owns one HIP stream. "encode" enqueues on that stream; completion is observed through
 * the read-back calls. One handle is used by one thread at a time (not internally locked). No C++
 * exception crosses this boundary: every call returns a status (>= 0 ok, < 0 an ISPH_E* code) and
 * isph_last_error() describes the last failure.
 */
#ifndef ISPH_H
#define ISPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-particle record, 48 bytes: `SphParticle` crates/engine/src/gpu_sph.rs:27-37 == WGSL `Particle`
 * shaders/sph_step.wgsl:45-49. SI units. */
typedef struct isph_particle {
    float pos[3];
    float h;
    float vel[3];
    float u;
    float mass;
    uint32_t mat;
    float rho;
    uint32_t prov;
} isph_particle;

/* Tillotson parameters, 48 bytes: `SphEos` gpu_sph.rs:42-55 == WGSL `Eos` sph_step.wgsl:50-54. */
typedef struct isph_eos {
    float rho0, a, b, cap_a, cap_b, e0, e_iv, e_cv, alpha, beta, _p0, _p1;
} isph_eos;

#ifdef __cplusplus
static_assert(sizeof(isph_particle) == 48, "SphParticle is 48 bytes (gpu_sph.rs:1520-1598)");
static_assert(sizeof(isph_eos) == 48, "SphEos is 48 bytes");
static_assert(offsetof(isph_particle, h) == 12 && offsetof(isph_particle, vel) == 16 &&
                  offsetof(isph_particle, u) == 28 && offsetof(isph_particle, mass) == 32 &&
                  offsetof(isph_particle, mat) == 36 && offsetof(isph_particle, rho) == 40 &&
                  offsetof(isph_particle, prov) == 44,
              "field offsets of the WGSL Particle struct");
#endif

typedef struct isph_ctx isph_t;

/* Status codes. */
#define ISPH_OK 0
#define ISPH_EAGAIN (-11) /* read-back not ready yet (take_readback's `None`) */
#define ISPH_EINVAL (-22)
#define ISPH_ENOMEM (-12)
#define ISPH_EDEVICE (-5) /* a HIP runtime call failed */
#define ISPH_ENODEV (-19)
#define ISPH_ERANGE (-34)

/* Gravity modes (isph_set_gravity). */
#define ISPH_GRAVITY_TREE 0   /* LBVH Barnes-Hut: bh_gravity.wgsl:122-394 */
#define ISPH_GRAVITY_DIRECT 1 /* exact softened direct sum: sph_step.wgsl:235-241 */
#define ISPH_GRAVITY_NONE 2   /* SPH forces only (parity probes) */

/* Precision modes (isph_set_precision). */
#define ISPH_PRECISION_FP32 0  /* plain: f32 state, f32 pair sums (the reference's GPU arithmetic) */
#define ISPH_PRECISION_MIXED 1 /* f64 position/velocity state; f64 SPH pair sums and per-particle rho, P, c; compensated f32 gravity sums */
/* The f64 reference of the mixed form's discrete step (same tree, walk decisions, pair sets and dt): f64
 * coordinates and velocities in every pair term, f64 node moments and f64 gravity; u stays f32. It exists to
 * measure what each form's rounding costs (a run minus the reference run); one system, leaf size 1, monopole
 * gravity with one particle per lane; no adjoint. */
#define ISPH_PRECISION_F64 2

/* Phase indices of isph_phase_times. */
#define ISPH_PHASE_SORT 0      /* bbox + Morton + radix sort + gather */
#define ISPH_PHASE_TREE 1      /* Karras build */
#define ISPH_PHASE_UPSWEEP 2   /* moments */
#define ISPH_PHASE_GRAVITY 3
#define ISPH_PHASE_DENSITY 4   /* density + per-particle EOS + Courant signal */
#define ISPH_PHASE_FORCE 5     /* SPH force + du/dt (+ gravity add, scatter) */
#define ISPH_PHASE_INTEGRATE 6 /* kick-drift, kick, dt */
#define ISPH_NUM_PHASES 7

/* ---- stepper (GpuSph, gpu_sph.rs:1086-1505) ------------------------------------------------------ */

/* GpuSph::new(device, capacity) gpu_sph.rs:1116. `device` is the HIP ordinal (the adapter choice of
 * gpu_host.rs:31). Capacity is clamped to >= 1 like the reference (:1117). NULL on failure. */
isph_t* isph_create(uint32_t capacity, int device);
void isph_destroy(isph_t* h);
/* Last error message of `h` (or, when h is NULL, of the calling thread's last isph_create or handle-less
 * call). */
const char* isph_last_error(const isph_t* h);

/* GpuSph::upload(queue, particles, eos, softening) gpu_sph.rs:1323-1340: particles beyond capacity are
 * dropped (gpu_store.rs:33), the EOS table is truncated to 16 entries (:1337). Returns the count kept.
 * Every particle's `mat` must index the table (ISPH_EINVAL otherwise). */
int isph_upload(isph_t* h, const isph_particle* particles, size_t n, const isph_eos* eos, size_t n_eos,
                float softening);
/* Replace the EOS table with f64 parameters (10 per material, eos.rs:27-47 field order). The f32 table
 * of isph_upload is otherwise widened exactly; this lets the per-particle FP64 EOS use the catalogue
 * values unrounded, as the f64 CPU path (`eos::Tillotson`) does. */
int isph_set_eos_f64(isph_t* h, const double* params10, size_t n_eos);
/* GpuSph::set_bulk(queue, center, r_core, vel, mass) gpu_sph.rs:1364-1380: the un-resolved bulk planet, a
 * Gauss gravity source (monopole outside r_core, linear inside) plus a non-injecting spherical floor;
 * r_core = 0 disables it. */
int isph_set_bulk(isph_t* h, double cx, double cy, double cz, double r_core, double vx, double vy, double vz,
                  double mass);
/* GpuSph::set_bulk_planar(queue, plane_point, up, g) gpu_sph.rs:1387-1403: uniform gravity g along -up and
 * a flat floor through plane_point (up normalised, +Y when zero). */
int isph_set_bulk_planar(isph_t* h, double px, double py, double pz, double ux, double uy, double uz, double g);
/* GpuSph::set_external_masses(queue, bodies) gpu_sph.rs:1412-1428: point masses (x, y, z, m per body) added
 * to every particle's gravity with the same softened form; at most 32 are kept, the count kept is returned. */
int isph_set_external_masses(isph_t* h, const double* pos_mass4, size_t n);
/* GpuSph::set_merge_budget(queue, budget) gpu_sph.rs:1432: merging is active while the particle count exceeds
 * `budget` (0 = never). */
int isph_set_merge_budget(isph_t* h, uint32_t budget);
/* GpuSph::encode_merge(enc) gpu_sph.rs:1439-1447: one merge pass (neighbours + density, pick, apply, retire,
 * sph_step.wgsl:374-541). */
int isph_encode_merge(isph_t* h);
/* GpuSph::encode_relax(enc, steps) gpu_sph.rs:1476-1481: `steps` damped relaxation steps (one force
 * evaluation each, then v = (v + (a + w^2 (x, y, 0)) dt) damp, x += v dt; u held) at the set_dt dt and damp. */
int isph_encode_relax(isph_t* h, uint32_t steps);
/* SphParams::omega gpu_sph.rs:70-76: the relaxation frame's spin about +z (rad/s), 0 = none. */
int isph_set_omega(isph_t* h, float omega);
/* GpuSph::count gpu_sph.rs:1449 */
uint32_t isph_count(const isph_t* h);
/* Ensemble upload (no reference counterpart: the batching form). `n_inst` independent systems share one
 * handle, one EOS table, softening and step controls: instance k owns particles [offsets[k], offsets[k+1])
 * (offsets[0] = 0, each instance non-empty, offsets[n_inst] <= capacity, n_inst <= 256). Every launch of a
 * step advances all instances; each has its own Morton box, its own subtree of the tree (walks never
 * cross instances), its own Courant dt and simulated time. Leaf size 1, tree or no gravity; the adjoint
 * and the pair potential energy take one system. Returns the particle count. */
int isph_upload_ensemble(isph_t* h, const isph_particle* particles, const uint32_t* offsets, uint32_t n_inst,
                         const isph_eos* eos, size_t n_eos, float softening);
/* Per-instance simulated time and last dt (up to `cap` instances); returns the instance count. */
int isph_ensemble_times(isph_t* h, double* t, float* dt, uint32_t cap);
/* GpuSph::set_dt(queue, dt, damp) gpu_sph.rs:1343 */
int isph_set_dt(isph_t* h, float dt, float damp);
/* GpuSph::set_av(queue, alpha, beta) gpu_sph.rs:1353 */
int isph_set_av(isph_t* h, float alpha, float beta);
/* Gravity configuration (the reference hard-codes THETA=0.5, BUCKET_K=1: gpu_gravity.rs:55,62).
 * leaf_size in [1, 64]; quadrupole != 0 adds the node quadrupole to accepted cells. */
int isph_set_gravity(isph_t* h, int mode, float theta, uint32_t leaf_size, int quadrupole);
/* Shape of the wave-cooperative tree walks (no reference counterpart; the autotuner's knobs). Particles are
 * walked in groups: the largest tree nodes whose box diagonal is <= group_extent * (largest h) and that
 * hold <= 64 particles (SPH walks, one per lane) or <= 64 * grav_particles_per_lane (gravity walk, 1 or 2
 * per lane). The SPH walks take a kept subtree of <= sph_expand particles whole as candidates.
 * Defaults: 4.0, 1, 4. Returns ISPH_EINVAL for group_extent <= 0, grav_particles_per_lane not 1 or 2, or
 * sph_expand outside [1, 64]. */
int isph_set_walk(isph_t* h, float group_extent, int grav_particles_per_lane, int sph_expand);
/* Candidate cache of the SPH walks (no reference counterpart). The density walk stores each walk group's
 * candidate batches and pair masks, in blocks of 32 chunks of 64 candidates (one block per walk wave to
 * start with, more from a pool), and the force pass of the same evaluation replays them; a group whose
 * list does not fit is walked again by the force pass with the same result bit for bit. This caps the
 * blocks in use: ISPH_CACHE_ALL (the default) = every allocated block, 0 = none (every group is walked
 * twice). It exists to exercise and time the walk fallback. */
#define ISPH_CACHE_ALL 0xffffffffu
int isph_set_candidate_cache(isph_t* h, uint32_t max_blocks);
int isph_set_precision(isph_t* h, int mode);
/* Reuse the tree, gravity and density of the previous evaluation when positions have not moved since
 * (eval 2 of step n and eval 1 of step n+1 see identical positions). Results are bit-identical. */
int isph_set_reuse(isph_t* h, int enable);
/* Graph mode (no reference counterpart): isph_encode_kdk / isph_step_cfl capture one KDK step (all its
 * launches) as a hipGraph on first use and replay it per step. Any configuration change or upload drops
 * the graph; the dt (encode) or cfl (step_cfl) value is part of its key. */
int isph_set_graph(isph_t* h, int enable);
/* Autotuner (no reference counterpart). Times `trial_steps` Courant steps per candidate from the current
 * state (restored after every trial) and keeps the fastest walk shape by coordinate descent over the leaf
 * size (one system, tree gravity: 1, 2, 4), the group extent (3, 4, 6), the gravity particles per lane
 * (1, 2) and the SPH whole-subtree size (1, 4, 8); the radix width is fixed (8-bit digits: one pass is
 * about 1.2 % of a 1M-particle evaluation, the most a 3-pass 10-bit sort could save). Results are
 * cached per size bucket (floor(log2 n)), ensemble flag, precision and gravity form in the text file
 * `cache_path` (may be NULL): a cached entry is applied without trials (returns 1), a tuned one is appended
 * (returns 0); an entry that isph_set_gravity / isph_set_walk would refuse for this handle is skipped. The
 * chosen shape is applied to the handle and written to `chosen` (may be NULL). */
typedef struct isph_tuning {
    uint32_t leaf_size;
    float group_extent;
    int grav_particles_per_lane;
    int sph_expand;
    double ms_per_step;
} isph_tuning;
int isph_autotune(isph_t* h, float cfl, uint32_t trial_steps, const char* cache_path, isph_tuning* chosen);

/* GpuSph::encode_kdk(enc, substeps) gpu_sph.rs:1484-1491: `substeps` KDK steps at the set_dt dt, each
 * = eval -> half-kick + drift -> eval -> half-kick (sph_step.wgsl:304-333). Enqueued, returns at once. */
int isph_encode_kdk(isph_t* h, uint32_t substeps);
/* The adaptive-dt KDK loop of tools/impact-run/src/main.rs:394-434 with the Courant reduction
 * (sph_step.wgsl:364-371, main.rs:345-356) done on the device: no host round trip per step. */
int isph_step_cfl(isph_t* h, float cfl, uint32_t steps);
/* Simulated time accumulated by isph_step_cfl / isph_encode_kdk (seconds, f64), and the last dt. */
int isph_sim_time(isph_t* h, double* t, float* last_dt);

/* GpuSph::begin_readback gpu_sph.rs:1497 / gpu_store.rs:126-153: no-op when empty or in flight. */
int isph_begin_readback(isph_t* h);
/* GpuSph::take_readback gpu_sph.rs:1502 / gpu_store.rs:156-170: the snapshot's particle count, or
 * ISPH_EAGAIN while pending or idle (the reference's `None`). Never blocks. */
int isph_take_readback(isph_t* h, isph_particle* out, size_t cap);
/* Blocking read-back for native callers (the offline driver's read_particles, main.rs:332-343). */
int isph_readback_sync(isph_t* h, isph_particle* out, size_t cap);
int isph_synchronize(isph_t* h);
/* GpuSph::particle_buffer (gpu_sph.rs:1452-1456): the device array of 48-byte particle records that the
 * renderer draws (shaders/sph_render.wgsl) without a copy. The state lives in per-field arrays on the
 * device, so this packs the current state into the handle's record array on its stream and waits; *dev_ptr
 * receives that array's device address and *count the particle count. The array belongs to the handle and
 * holds this snapshot until the next isph_particle_buffer or read-back call. */
int isph_particle_buffer(isph_t* h, void** dev_ptr, uint32_t* count);
/* The handle's HIP stream (a hipStream_t), for interop callers that order their own device work after the
 * handle's. */
void* isph_stream(isph_t* h);

/* One force evaluation at the current state (the reference's force_eval gpu_sph.rs:1468-1474:
 * density + forces). Writes, in upload order: acc4 = (ax, ay, az, du/dt) per particle, rho. Either
 * output may be NULL. */
int isph_force_eval(isph_t* h, float* acc4, float* rho);

/* ---- gravity (GpuGravity, gpu_gravity.rs:79-561) --------------------------------------------------- */
/* GpuGravity::accelerations (mode ISPH_GRAVITY_DIRECT, :187) / accelerations_tree (mode
 * ISPH_GRAVITY_TREE, :207, theta and leaf size from isph_set_gravity): f64 in, f32 on the device, f64
 * out. n must not exceed the handle's capacity. Does not touch the stepper's particle state. */
int isph_gravity_accelerations(isph_t* h, const double* pos3, const double* mass, size_t n, double softening,
                               double* acc3, int mode);

/* No counterpart in the reference: the direct mode of isph_gravity_accelerations for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Takes no handle; pos3[3n], mass[n] and acc3[3n] are DEVICE arrays of the current device,
 * f64 in and out, f32 compensated sums on the device (bodies (x, y, z, m) rounded to f32 as the allocating
 * call does). Enqueued on `stream` (a hipStream_t, NULL = the null stream); returns without waiting and
 * calls no allocator, memset or synchronisation, so it can be captured. Its staging lives in a fixed
 * per-device scratch array of 32 MiB held until the process ends: n <= 1,048,576 (ISPH_ERANGE above);
 * calls that overlap on one GPU share that array and race. Errors: isph_last_error(NULL). */
int isph_gravity_direct_noalloc(const double* pos3, const double* mass, size_t n, double softening, double* acc3,
                                void* stream);

/* ---- diagnostics ---------------------------------------------------------------------------------- */
/* Totals in f64 over the live particles: out[0]=KE, [1]=IE, [2]=softened PE (exact O(N^2) device sum
 * when with_pe != 0, else 0), [3..5]=angular momentum about the COM, [6]=total mass.
 * gpu_sph.rs:940-961, impact-run main.rs:452-472, 528-536. */
int isph_diagnostics(isph_t* h, int with_pe, double* out7);
/* Accumulated per-phase milliseconds (ISPH_PHASE_*) since the last reset, when timing is enabled. */
int isph_set_timing(isph_t* h, int enable);
int isph_phase_times(isph_t* h, double* ms, size_t n);
/* Gravity-kernel statistics since the last reset: out[0] = launches, [1] = total kernel ms (HIP events on
 * the handle's stream), [2] = monopole cell interactions, [3] = particle-particle interactions,
 * [4] = quadrupole cell interactions, [5] = density pairs tested, [6] = force pairs tested,
 * [7] = algorithmic gravity flops (interactions times the per-interaction flop constants). */
int isph_kernel_stats(isph_t* h, double* out8);
/* Tree-walk statistics since the last reset: out[0..2] = gravity walk rounds, nodes popped, deepest stack;
 * out[3..5] the same for the density walk; out[6..8] for the force walk. */
int isph_walk_stats(isph_t* h, double* out9);
int isph_reset_stats(isph_t* h);
/* Energy (J, f64) that the u >= 0 clamp of the half-kicks (sph_step.wgsl:310, hydrostatic.rs:457) has added
 * since the last isph_reset_stats: the sum of m (-u_new) over every clamped update. No reference
 * counterpart; it separates the clamp from the other sources of energy drift. */
int isph_clamp_energy(isph_t* h, double* joules);

/* ---- adjoint (no reference counterpart: the differentiable form of the step) ------------------------ */
/* Smoothed disk mass: loss = sum_i m_i sig(k_r (r_i / radius - r_cut)) sig(k_j (j_i / sqrt(G M r_i) - 1)),
 * r_i = |x_i - X|, j_i = axis . ((x_i - X) x (v_i - V)), X, V, M the centre of mass, its velocity and the
 * total mass; sig(z) = 1 / (1 + e^-z). SI units, f64. */
typedef struct isph_loss {
    double radius; /* length unit of the cut (the target radius), m */
    double r_cut;  /* cut radius in units of `radius` */
    double k_r;    /* sharpness of the radial sigmoid */
    double k_j;    /* sharpness of the angular-momentum sigmoid */
    double axis[3];
} isph_loss;

/* Loss of the current state (mixed form). */
int isph_loss_value(isph_t* h, const isph_loss* loss, double* value);
/* Replace the f64 position / velocity state exactly (either f64 form; u, h, mass unchanged). */
int isph_set_state_f64(isph_t* h, const double* pos3, const double* vel3);
/* Read the f64 position / velocity state (either f64 form), upload order; either output may be NULL. */
int isph_state_f64(isph_t* h, double* pos3, double* vel3);
/* Reverse-mode gradient of `loss` after `steps` KDK steps from the current state (mixed form; gravity direct,
 * none, or the monopole tree walk with one particle per lane). Time steps: dts_in[steps] if not NULL, else
 * the device Courant dt at `cfl`; they are held fixed in the derivative. The forward run stores a state
 * every `checkpoint_every` steps; the reverse pass recomputes each segment from its checkpoint and walks it
 * back one evaluation at a time with the tree topology and walk decisions of the forward evaluation.
 * Outputs (any may be NULL): the loss; d loss / d (position[3n], velocity[3n], u[n], mass[n]) of the start
 * state in upload order; d loss / d (a, b, A, B) of each material's Tillotson parameters [16][4]; the dt
 * sequence [steps]; ms2 = {forward ms, reverse ms} from HIP events on the handle's stream. The handle is
 * left at the forward run's final state. */
int isph_vjp_run(isph_t* h, float cfl, const float* dts_in, uint32_t steps, uint32_t checkpoint_every,
                 const isph_loss* loss, double* loss_out, double* dpos3, double* dvel3, double* du, double* dmass,
                 double* deos4, float* dts_out, double* ms2);

/* ---- measurement (DiskView, gpu_sph.rs:652-823; accretion::find_clumps, accretion.rs:158-301) --------- */
/* The disk of one system: DiskView's fields (gpu_sph.rs:661-686) plus impact-run's split of the remnant and
 * escaped mass by source body and its clump counts (tools/impact-run/src/main.rs:732-778). f64, SI. */
typedef struct isph_disk {
    int32_t valid;     /* 0 where DiskView::of is None (fewer than 2 live particles, or no mass): rest unset */
    uint32_t n_live;   /* particles with positive finite mass and finite position and velocity */
    uint32_t n_disk;   /* disk members: bound to the remnant with perigee above its radius */
    uint32_t n_clumps; /* disk clumps with at least 2 members (impact-run's n_clumps) */
    uint32_t n_bound;  /* of those, self-bound (impact-run's n_bound) */
    uint32_t _pad;
    double com[3], v_com[3];                 /* frame: centre of mass of the live field and its velocity */
    double r_remnant, m_remnant, mu;         /* smallest radius enclosing 85 % of the mass, that mass, G m */
    double disk_prov0, disk_other, escaped;  /* disk mass from source body 0 / the rest; unbound mass */
    double mean_h;                           /* mean smoothing length of the live field */
    double remnant_prov0, remnant_other;     /* bound with perigee inside the remnant, by source body */
    double escaped_prov0;                    /* unbound mass from source body 0 */
    double moon_mass;                        /* largest clump that accretes (disk_stats_json's "moon"), or 0 */
} isph_disk;

/* One friends-of-friends clump: `Clump` accretion.rs:37-51 (members are returned separately).
 * Clump::accretes() == n_members >= 2 && bound && outside_roche. */
typedef struct isph_clump {
    uint32_t id;        /* smallest member index: a position in isph_disk_members' list, or an input index */
    uint32_t n_members;
    int32_t bound;         /* internal_ke + self_pe < 0 */
    int32_t outside_roche; /* COM beyond the remnant's fluid Roche limit for this clump's density */
    double mass, com_pos[3], com_vel[3], rho, radius, internal_ke, self_pe, ang_mom[3], thermal_ke;
} isph_clump;

#ifdef __cplusplus
static_assert(sizeof(isph_disk) == 160, "isph_disk layout");
static_assert(sizeof(isph_clump) == 136, "isph_clump layout");
#endif

/* DiskView::clumps' softening (1e4 m) and linking length factor (2 x mean h), gpu_sph.rs:796-808. */
int isph_set_measure(isph_t* h, double clump_softening, double link_factor);
/* DiskView::of (gpu_sph.rs:689-785) of the current state, on the device, for every instance (one record for
 * a single system), with the disk's clumps (DiskView::clumps, gpu_sph.rs:793-809). Writes min(cap, count)
 * records and returns the instance count. The members and clumps below refer to this measurement until the
 * next isph_disk_view or isph_find_clumps call. */
int isph_disk_view(isph_t* h, isph_disk* out, uint32_t cap);
/* DiskView's `idx` (indices into the instance's own particles, ascending) and each member's clump id.
 * Either output may be NULL. Returns the member count. */
int isph_disk_members(isph_t* h, uint32_t inst, uint32_t* idx, uint32_t* clump_id, uint32_t cap);
/* The instance's clumps with at least min_members members, heaviest first (ties by id); an instance with
 * fewer than 2 disk members has none (gpu_sph.rs:794). moonlets(k) = the ones that accrete with
 * n_members >= k (:813-820). Returns the count; writes min(cap, count). */
int isph_disk_clumps(isph_t* h, uint32_t inst, uint32_t min_members, isph_clump* out, uint32_t cap);
/* disk_stats_json (gpu_sph.rs:825-851) of the measured instance, snprintf-style: returns the length of the
 * full text ("null" when the view is None). */
int isph_disk_stats_json(isph_t* h, uint32_t inst, char* buf, size_t cap);
/* accretion::find_clumps (accretion.rs:158-301) on caller arrays (f64, n <= capacity): friends-of-friends
 * within linking_length, then per clump the statistics, the softened self potential energy (an exact pair
 * sum at every size) times g / G, and the Roche gate against the given central body. Writes the clumps
 * with >= min_members members, heaviest first, and labels[n] = each particle's clump id (the smallest
 * member index). Returns the count. Leaves no disk view behind. */
int isph_find_clumps(isph_t* h, const double* pos3, const double* vel3, const double* mass, const double* rho,
                     size_t n, double linking_length, double g, double softening, const double* central_pos3,
                     double central_mass, double central_radius, uint32_t min_members, isph_clump* out,
                     uint32_t cap, uint32_t* labels);

/* ---- stage probes (parity of the bit-exact stages against the oracle) ------------------------------- */
/* bbox + Morton of the current positions: codes in upload order, lo3/hi3 the f32 bbox (cs_bbox/cs_morton,
 * bh_gravity.wgsl:67-108). */
int isph_probe_morton(isph_t* h, uint32_t* codes, float* lo3, float* hi3);
/* Sorted (code, index) pairs (the CPU sort of gpu_gravity.rs:378-380, done by the device radix sort). */
int isph_probe_sort(isph_t* h, uint32_t* sorted_codes, uint32_t* order);
/* Karras tree and moments over L = ceil(n / leaf_size) leaves (bh_gravity.wgsl:122-241): left/right
 * [L-1], parent [2L-1], com4 [2L-1][4] (xyz, mass), bmin4/bmax4 [2L-1][4] (xyz, w = max h / MAC radius).
 * Returns L. Any output may be NULL. */
int isph_probe_tree(isph_t* h, uint32_t* left, uint32_t* right, uint32_t* parent, float* com4, float* bmin4,
                    float* bmax4);

#ifdef __cplusplus
}
#endif
#endif /* ISPH_H */
