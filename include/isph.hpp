/*
 * isph.hpp — C++17 mirror of integrity-engine's GPU stepper and gravity types over the C-ABI of isph.h.
 *
 * Names, argument order and return types follow the Rust surface (crates/engine/src/gpu_sph.rs:1115-1505,
 * gpu_gravity.rs:79-640); the wgpu `&Device` / `&Queue` / `&mut CommandEncoder` arguments are gone, since the
 * handle owns its device and stream. `Option<T>` is std::optional<T>, `Vec<DVec3>` std::vector<DVec3>.
 * The reference reports no errors from these types (silent clamps, returned counts, panics on device loss);
 * here a failing call throws isph_engine::Error carrying the ISPH_E* status and the library's message, and
 * the clamps are the library's (capacity on upload, 16 EOS entries, 32 external masses).
 *
 * Header only: include it and link libisph.so. Not thread-safe per object, like the Rust `&mut self`.
 */
#ifndef ISPH_HPP
#define ISPH_HPP

#include <algorithm>
#include <array>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "isph.h"

namespace isph_engine {

using SphParticle = isph_particle;  // gpu_sph.rs:27-37
using SphEos = isph_eos;            // gpu_sph.rs:42-55
using DVec3 = std::array<double, 3>;

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& message) : std::runtime_error(message), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

namespace detail {
inline int check(const isph_t* h, int rc) {
    if (rc < 0 && rc != ISPH_EAGAIN) throw Error(rc, isph_last_error(h));
    return rc;
}

// One owned handle: move-only, destroyed with its owner.
class Handle {
public:
    Handle() = default;
    Handle(uint32_t capacity, int device) : h_(isph_create(capacity, device)) {
        if (!h_) throw Error(ISPH_ENODEV, std::string("isph_create: ") + isph_last_error(nullptr));
    }
    Handle(Handle&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    Handle& operator=(Handle&& o) noexcept {
        if (this != &o) {
            reset();
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    ~Handle() { reset(); }
    void reset() {
        if (h_) isph_destroy(h_);
        h_ = nullptr;
    }
    isph_t* get() const { return h_; }
    explicit operator bool() const { return h_ != nullptr; }

private:
    isph_t* h_ = nullptr;
};
}  // namespace detail

/* The lent particle array (gpu_sph.rs:1452-1456, `particle_buffer`): device address and count. */
struct ParticleBuffer {
    void* device_ptr;
    uint32_t count;
};

/* One force evaluation in upload order (isph_force_eval): acceleration, du/dt, density. */
struct ForceEval {
    std::vector<std::array<float, 3>> acc;
    std::vector<float> dudt;
    std::vector<float> rho;
};

/* GpuSph (gpu_sph.rs:1086-1505). */
class GpuSph {
public:
    /* GpuSph::new(device, capacity) :1116; `device` is the HIP ordinal. */
    explicit GpuSph(uint32_t capacity, int device = 0) : h_(capacity, device) {}

    /* upload(queue, particles, eos, softening) :1323-1340. Returns the count kept (the reference drops the
     * excess silently and returns ()). */
    uint32_t upload(const std::vector<SphParticle>& particles, const std::vector<SphEos>& eos, float softening) {
        return (uint32_t)detail::check(h(), isph_upload(h(), particles.data(), particles.size(), eos.data(),
                                                         eos.size(), softening));
    }
    /* set_dt(queue, dt, damp) :1343 */
    void set_dt(float dt, float damp) { detail::check(h(), isph_set_dt(h(), dt, damp)); }
    /* set_av(queue, alpha, beta) :1353 */
    void set_av(float alpha, float beta) { detail::check(h(), isph_set_av(h(), alpha, beta)); }
    /* set_bulk(queue, center, r_core, vel, mass) :1364-1380 */
    void set_bulk(const DVec3& center, double r_core, const DVec3& vel, double mass) {
        detail::check(h(), isph_set_bulk(h(), center[0], center[1], center[2], r_core, vel[0], vel[1], vel[2], mass));
    }
    /* set_bulk_planar(queue, plane_point, up, g) :1387-1403 */
    void set_bulk_planar(const DVec3& plane_point, const DVec3& up, double g) {
        detail::check(h(), isph_set_bulk_planar(h(), plane_point[0], plane_point[1], plane_point[2], up[0], up[1],
                                                up[2], g));
    }
    /* set_external_masses(queue, bodies) -> usize :1412-1428: the count kept (at most 32). */
    size_t set_external_masses(const std::vector<std::pair<DVec3, double>>& bodies) {
        std::vector<double> pm;
        pm.reserve(4 * bodies.size());
        for (const auto& b : bodies) pm.insert(pm.end(), {b.first[0], b.first[1], b.first[2], b.second});
        return (size_t)detail::check(h(), isph_set_external_masses(h(), pm.data(), bodies.size()));
    }
    /* set_merge_budget(queue, budget) :1432 */
    void set_merge_budget(uint32_t budget) { detail::check(h(), isph_set_merge_budget(h(), budget)); }
    /* encode_merge(enc) :1439-1447 */
    void encode_merge() { detail::check(h(), isph_encode_merge(h())); }
    /* count() :1449 */
    uint32_t count() const { return isph_count(h()); }
    /* particle_buffer() :1452-1456 */
    ParticleBuffer particle_buffer() {
        ParticleBuffer b{nullptr, 0};
        detail::check(h(), isph_particle_buffer(h(), &b.device_ptr, &b.count));
        return b;
    }
    /* encode_relax(enc, steps) :1476-1481 */
    void encode_relax(uint32_t steps) { detail::check(h(), isph_encode_relax(h(), steps)); }
    /* encode_kdk(enc, substeps) :1484-1491 */
    void encode_kdk(uint32_t substeps) { detail::check(h(), isph_encode_kdk(h(), substeps)); }
    /* begin_readback(device, queue) :1497 */
    void begin_readback() { detail::check(h(), isph_begin_readback(h())); }
    /* take_readback() -> Option<Vec<SphParticle>> :1502: nullopt while pending or idle. Never blocks. */
    std::optional<std::vector<SphParticle>> take_readback() {
        std::vector<SphParticle> out(count());
        const int rc = detail::check(h(), isph_take_readback(h(), out.data(), out.size()));
        if (rc == ISPH_EAGAIN) return std::nullopt;
        out.resize(std::min(out.size(), (size_t)rc));  // an upload in between discards the snapshot
        return out;
    }

    // ---- beyond the reference surface (isph.h) ----
    void set_eos_f64(const std::vector<std::array<double, 10>>& params) {
        detail::check(h(), isph_set_eos_f64(h(), params.empty() ? nullptr : params[0].data(), params.size()));
    }
    void set_gravity(int mode, float theta = 0.5f, uint32_t leaf_size = 1, bool quadrupole = false) {
        detail::check(h(), isph_set_gravity(h(), mode, theta, leaf_size, quadrupole ? 1 : 0));
    }
    void set_precision(int mode) { detail::check(h(), isph_set_precision(h(), mode)); }
    void set_walk(float group_extent, int grav_particles_per_lane, int sph_expand) {
        detail::check(h(), isph_set_walk(h(), group_extent, grav_particles_per_lane, sph_expand));
    }
    void set_omega(float omega) { detail::check(h(), isph_set_omega(h(), omega)); }
    /* The adaptive-dt loop of tools/impact-run/src/main.rs:394-434 with the Courant minimum on the device. */
    void step_cfl(float cfl, uint32_t steps) { detail::check(h(), isph_step_cfl(h(), cfl, steps)); }
    /* Simulated time (s) and the last dt. */
    std::pair<double, float> sim_time() {
        double t = 0.0;
        float dt = 0.0f;
        detail::check(h(), isph_sim_time(h(), &t, &dt));
        return {t, dt};
    }
    std::vector<SphParticle> readback_sync() {
        std::vector<SphParticle> out(count());
        out.resize(std::min(out.size(), (size_t)detail::check(h(), isph_readback_sync(h(), out.data(), out.size()))));
        return out;
    }
    void synchronize() { detail::check(h(), isph_synchronize(h())); }
    ForceEval force_eval() {
        const size_t n = count();
        std::vector<float> acc4(4 * n);
        ForceEval e;
        e.rho.resize(n);
        detail::check(h(), isph_force_eval(h(), acc4.data(), e.rho.data()));
        e.acc.resize(n);
        e.dudt.resize(n);
        for (size_t i = 0; i < n; ++i) {
            e.acc[i] = {acc4[4 * i], acc4[4 * i + 1], acc4[4 * i + 2]};
            e.dudt[i] = acc4[4 * i + 3];
        }
        return e;
    }
    /* KE, IE, PE (exact device sum when with_pe), angular momentum about the COM, total mass. */
    std::array<double, 7> diagnostics(bool with_pe) {
        std::array<double, 7> out{};
        detail::check(h(), isph_diagnostics(h(), with_pe ? 1 : 0, out.data()));
        return out;
    }
    /* The handle's hipStream_t, for callers that order their own device work against it. */
    void* stream() const { return isph_stream(h()); }
    isph_t* handle() const { return h_.get(); }

    // ---- batching: ensembles, captured steps, autotuner ----
    /* Independent systems in one handle (isph_upload_ensemble); returns the offsets of each instance. */
    std::vector<uint32_t> upload_ensemble(const std::vector<std::vector<SphParticle>>& instances,
                                          const std::vector<SphEos>& eos, float softening) {
        std::vector<uint32_t> off{0u};
        std::vector<SphParticle> all;
        for (const auto& r : instances) {
            all.insert(all.end(), r.begin(), r.end());
            off.push_back((uint32_t)all.size());
        }
        detail::check(h(), isph_upload_ensemble(h(), all.data(), off.data(), (uint32_t)instances.size(),
                                                eos.data(), eos.size(), softening));
        return off;
    }
    void set_graph(bool enable) { detail::check(h(), isph_set_graph(h(), enable ? 1 : 0)); }
    void set_reuse(bool enable) { detail::check(h(), isph_set_reuse(h(), enable ? 1 : 0)); }
    void set_candidate_cache(uint32_t max_blocks) { detail::check(h(), isph_set_candidate_cache(h(), max_blocks)); }
    /* Returns (taken from the cache file, the shape applied). */
    std::pair<bool, isph_tuning> autotune(float cfl, uint32_t trial_steps, const char* cache_path = nullptr) {
        isph_tuning t{};
        const int rc = detail::check(h(), isph_autotune(h(), cfl, trial_steps, cache_path, &t));
        return {rc == 1, t};
    }

    // ---- measurement (DiskView gpu_sph.rs:652-851, accretion::find_clumps accretion.rs:158-301) ----
    void set_measure(double clump_softening, double link_factor) {
        detail::check(h(), isph_set_measure(h(), clump_softening, link_factor));
    }
    /* DiskView::of of every instance (one record for a single system); `valid` == 0 is the reference's None. */
    std::vector<isph_disk> disk_view() {
        std::vector<isph_disk> out(256);
        out.resize((size_t)detail::check(h(), isph_disk_view(h(), out.data(), (uint32_t)out.size())));
        return out;
    }
    /* DiskView::clumps (min_members 1) and moonlets(k) (the accreting ones with >= k members). */
    std::vector<isph_clump> disk_clumps(uint32_t inst = 0, uint32_t min_members = 1) {
        const int n = detail::check(h(), isph_disk_clumps(h(), inst, min_members, nullptr, 0));
        std::vector<isph_clump> out((size_t)n);
        detail::check(h(), isph_disk_clumps(h(), inst, min_members, out.data(), (uint32_t)out.size()));
        return out;
    }
    std::vector<isph_clump> moonlets(uint32_t k, uint32_t inst = 0) {
        std::vector<isph_clump> out;
        for (const auto& c : disk_clumps(inst, k))
            if (c.n_members >= 2 && c.bound && c.outside_roche) out.push_back(c);
        return out;
    }
    /* DiskView's member indices and each member's clump id. */
    std::pair<std::vector<uint32_t>, std::vector<uint32_t>> disk_members(uint32_t inst = 0) {
        const int n = detail::check(h(), isph_disk_members(h(), inst, nullptr, nullptr, 0));
        std::vector<uint32_t> idx((size_t)n), clump((size_t)n);
        detail::check(h(), isph_disk_members(h(), inst, idx.data(), clump.data(), (uint32_t)n));
        return {idx, clump};
    }
    /* disk_stats_json (gpu_sph.rs:825-851). */
    std::string disk_stats_json(uint32_t inst = 0) {
        const int n = detail::check(h(), isph_disk_stats_json(h(), inst, nullptr, 0));
        std::string buf((size_t)n + 1, '\0');
        detail::check(h(), isph_disk_stats_json(h(), inst, &buf[0], buf.size()));
        buf.resize((size_t)n);
        return buf;
    }

    /* accretion::find_clumps (accretion.rs:158-301) on caller arrays (n <= capacity): the clumps with at least
     * min_members members, heaviest first, and every particle's clump id (its smallest member index). */
    struct Clumps {
        std::vector<isph_clump> clumps;
        std::vector<uint32_t> labels;
    };
    Clumps find_clumps(const std::vector<DVec3>& pos, const std::vector<DVec3>& vel, const std::vector<double>& mass,
                       const std::vector<double>& rho, double linking_length, double g, double softening,
                       const DVec3& central_pos, double central_mass, double central_radius,
                       uint32_t min_members) {
        const size_t n = pos.size();
        if (vel.size() != n || mass.size() != n || rho.size() != n)
            throw Error(ISPH_EINVAL, "find_clumps: arrays differ in length");
        Clumps r;
        r.clumps.resize(n / std::max<uint32_t>(min_members, 1u));  // no more clumps than that
        r.labels.resize(n);
        const int k = detail::check(
            h(), isph_find_clumps(h(), n ? pos[0].data() : nullptr, n ? vel[0].data() : nullptr, mass.data(),
                                  rho.data(), n, linking_length, g, softening, central_pos.data(), central_mass,
                                  central_radius, min_members, r.clumps.data(), (uint32_t)r.clumps.size(),
                                  r.labels.data()));
        r.clumps.resize(std::min(r.clumps.size(), (size_t)k));
        return r;
    }

    // ---- differentiation (mixed form) ----
    void set_state_f64(const std::vector<DVec3>& pos, const std::vector<DVec3>& vel) {
        if (pos.size() != count() || vel.size() != count()) throw Error(ISPH_EINVAL, "state size differs from count()");
        detail::check(h(), isph_set_state_f64(h(), pos.empty() ? nullptr : pos[0].data(),
                                              vel.empty() ? nullptr : vel[0].data()));
    }
    double loss_value(const isph_loss& loss) {
        double v = 0.0;
        detail::check(h(), isph_loss_value(h(), &loss, &v));
        return v;
    }
    struct Vjp {
        double loss = 0.0;
        std::vector<DVec3> dpos, dvel;
        std::vector<double> du, dmass;
        std::vector<std::array<double, 4>> deos;  // d loss / d (a, b, A, B) per material
        std::vector<float> dts;
        double forward_ms = 0.0, reverse_ms = 0.0;
    };
    /* Reverse-mode gradient of `loss` after `steps` KDK steps (isph_vjp_run); dts empty = Courant at cfl. */
    Vjp vjp_run(const isph_loss& loss, uint32_t steps, uint32_t checkpoint_every, float cfl,
                const std::vector<float>& dts = {}) {
        if (!dts.empty() && dts.size() != steps) throw Error(ISPH_EINVAL, "dts must hold one dt per step");
        const size_t n = count();
        Vjp r;
        r.dpos.resize(n);
        r.dvel.resize(n);
        r.du.resize(n);
        r.dmass.resize(n);
        r.deos.resize(16);
        r.dts.resize(steps);
        double ms[2] = {0.0, 0.0};
        detail::check(h(), isph_vjp_run(h(), cfl, dts.empty() ? nullptr : dts.data(), steps, checkpoint_every, &loss,
                                        &r.loss, n ? r.dpos[0].data() : nullptr, n ? r.dvel[0].data() : nullptr,
                                        r.du.data(), r.dmass.data(), r.deos[0].data(), r.dts.data(), ms));
        r.forward_ms = ms[0];
        r.reverse_ms = ms[1];
        return r;
    }

    // ---- timing and statistics ----
    void set_timing(bool enable) { detail::check(h(), isph_set_timing(h(), enable ? 1 : 0)); }
    std::array<double, ISPH_NUM_PHASES> phase_times() {
        std::array<double, ISPH_NUM_PHASES> ms{};
        detail::check(h(), isph_phase_times(h(), ms.data(), ms.size()));
        return ms;
    }
    void reset_stats() { detail::check(h(), isph_reset_stats(h())); }
    double clamp_energy() {
        double j = 0.0;
        detail::check(h(), isph_clamp_energy(h(), &j));
        return j;
    }

private:
    isph_t* h() const { return h_.get(); }
    detail::Handle h_;
};

/* GpuGravity (gpu_gravity.rs:79-230): the direct sum and the LBVH tree on explicit f64 positions and masses.
 * A handle of sufficient capacity is created on first use and grown when a larger set arrives. */
class GpuGravity {
public:
    /* GpuGravity::new(device) :94 */
    explicit GpuGravity(int device = 0) : device_(device) {}
    /* accelerations(device, queue, pos, mass, softening) -> Vec<DVec3> :187: the exact softened sum. */
    std::vector<DVec3> accelerations(const std::vector<DVec3>& pos, const std::vector<double>& mass,
                                     double softening) {
        return run(pos, mass, softening, ISPH_GRAVITY_DIRECT);
    }
    /* accelerations_tree(device, queue, pos, mass, softening) -> Vec<DVec3> :207: LBVH Barnes-Hut, theta
     * 0.5, one particle per leaf (THETA :55, BUCKET_K :61). */
    std::vector<DVec3> accelerations_tree(const std::vector<DVec3>& pos, const std::vector<double>& mass,
                                          double softening) {
        return run(pos, mass, softening, ISPH_GRAVITY_TREE);
    }

private:
    std::vector<DVec3> run(const std::vector<DVec3>& pos, const std::vector<double>& mass, double softening,
                           int mode) {
        if (pos.size() != mass.size()) throw Error(ISPH_EINVAL, "positions and masses differ in length");
        const size_t n = pos.size();
        std::vector<DVec3> acc(n);
        if (n == 0) return acc;
        if (!h_ || n > cap_) {
            h_ = detail::Handle((uint32_t)n, device_);
            cap_ = n;
            detail::check(h_.get(), isph_set_gravity(h_.get(), ISPH_GRAVITY_TREE, 0.5f, 1, 0));
        }
        detail::check(h_.get(), isph_gravity_accelerations(h_.get(), pos[0].data(), mass.data(), n, softening,
                                                           acc[0].data(), mode));
        return acc;
    }
    int device_;
    size_t cap_ = 0;
    detail::Handle h_;
};

/* GravityField (gpu_gravity.rs:583-640), native form: the field is computed and read back in the call, so
 * it is never deferred and never `None`. Below TREE_KNEE (:50) particles the exact sum runs, at or above it
 * the tree, as the reference's dispatch does. */
class GravityField {
public:
    static constexpr size_t kTreeKnee = 24000;  // TREE_KNEE, gpu_gravity.rs:50
    explicit GravityField(int device = 0) : pipeline_(device) {}
    /* accelerations(pos, mass, softening) -> Option<Vec<DVec3>> :607 */
    std::optional<std::vector<DVec3>> accelerations(const std::vector<DVec3>& pos, const std::vector<double>& mass,
                                                    double softening) {
        if (pos.empty()) return std::vector<DVec3>{};
        return pos.size() < kTreeKnee ? pipeline_.accelerations(pos, mass, softening)
                                      : pipeline_.accelerations_tree(pos, mass, softening);
    }

private:
    GpuGravity pipeline_;
};

}  // namespace isph_engine

#endif  // ISPH_HPP
