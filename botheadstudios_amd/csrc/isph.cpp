// C-ABI host side of the MI355X particle step: handle lifetime, device memory, the per-evaluation
// pipeline and the KDK drivers. Mirrors GpuSph (crates/engine/src/gpu_sph.rs:1086-1505) and GpuGravity
// (crates/engine/src/gpu_gravity.rs:79-561); see include/isph.h for the per-entry-point mapping.
#include "../../include/isph.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "adjoint.h"
#include "measure.h"
#include "state.h"

using namespace isph;

namespace {

thread_local std::string g_create_error;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_OK(call)                                                                                        \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            throw HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " (" + std::to_string(e_) + ")"); \
    } while (0)

}  // namespace

namespace isph {
void set_thread_error(const std::string& msg) { g_create_error = msg; }
}  // namespace isph

namespace {

struct EvPair {
    int phase;
    hipEvent_t a, b;
};

}  // namespace

struct isph_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Dev d;
    StepParams p;
    int gravity_mode = ISPH_GRAVITY_TREE;
    int precision = ISPH_PRECISION_FP32;
    float softening = 0.0f;
    float dt_set = 0.0f;
    float damp = 1.0f;
    bool reuse_enabled = false;
    bool reuse_valid = false;
    uint32_t n_eos = 0;
    std::string err;

    // read-back
    isph_particle* pinned = nullptr;
    hipEvent_t rb_event = nullptr;
    bool rb_inflight = false;
    uint32_t rb_count = 0;

    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<EvPair> ev_pairs;
    double phase_ms[ISPH_NUM_PHASES] = {0};
    double grav_launches = 0;

    // gravity-API scratch
    float4* gpos = nullptr;
    float* gmass = nullptr;

    // adjoint arrays (allocated on first use)
    Adj adj;

    // captured step (isph_set_graph): one KDK step as a hipGraph, replayed per step. Any change of the
    // configuration or the particle set drops it; cfl / dt values are part of its key.
    bool graph_on = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_mode = -1;  // 0 fixed dt, 1 Courant dt
    float graph_value = 0.0f;

    // measurement (isph_disk_view / isph_find_clumps): device buffers allocated on first use, the instance
    // offsets of the current upload, and the last measurement's host copies
    Meas meas;
    std::vector<uint32_t> hioff{0u, 0u};
    double clump_soft = 1.0e4;  // gpu_sph.rs:805
    double link_factor = 2.0;   // gpu_sph.rs:803
    std::vector<double> mview;
    std::vector<uint32_t> mmoff;
    std::vector<unsigned long long> mcsum;
};

namespace {

// one slot per 256-particle block of the kick kernels (their clamped-energy sums)
size_t inject_slots(const Dev& d) { return ((size_t)d.cap + 255) / 256; }

template <class T>
void dalloc(T** p, size_t count) {
    HIP_OK(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
}
// For the buffer sets allocated on first use: a member that an interrupted earlier call (out of memory)
// already allocated is kept. Each set is allocated with the member its ensure_* function tests last, so a
// set is either complete or retried, never taken as complete with members missing.
template <class T>
void dalloc_once(T** p, size_t count) {
    if (!*p) dalloc(p, count);
}

// Host-to-device transfers and fills that the handle's kernels read are ordered on the handle's stream. That
// stream is non-blocking, so it is not ordered after work on the legacy null stream, and a synchronous
// hipMemcpy from pageable host memory may return once the data is staged, before its DMA has landed: a kernel
// launched next on the stream could read the old contents of the destination (for an upload, material ids
// that index the EOS table among them).
void h2d(hipStream_t s, void* dst, const void* src, size_t bytes) {
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));  // the host buffer may be a temporary
}
void fill0(hipStream_t s, void* dst, size_t bytes) { HIP_OK(hipMemsetAsync(dst, 0, bytes, s)); }

void free_all(isph_ctx* c) {
    Dev& d = c->d;
    void* ptrs[] = {d.posh, d.velu, d.mass, d.mat, d.prov, d.rho, d.acc, d.pos64, d.vel64, d.keys, d.vals,
                    d.keys_alt, d.vals_alt, d.hist, d.spos, d.svel, d.sh, d.smat, d.srho, d.spr, d.scs, d.sgrav,
                    d.lcode, d.left, d.right, d.parent, d.cslot, d.nrange, d.nlevel, d.carry, d.ncarry, d.groups, d.ggroups, d.chead,
                    d.cnext, d.ccnt, d.cpool, d.cmask, d.gflag, d.gscan, d.srho64, d.spr64, d.scs64, d.acc64, d.nG, d.nS, d.nA, d.nQ0, d.nQ1, d.scal, d.dt,
                    d.time, d.counters, d.red, d.inject, d.eos, d.aos, c->gpos, c->gmass};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    void* eptrs[] = {d.inst, d.mcode, d.lkey, d.ioff, d.ibox, d.iroot, d.isig, d.idt, d.itime, d.ext, d.mtarget,
                     d.slo, d.svlo, d.ndepth, d.nA64, d.sgrav64};
    for (void* q : eptrs)
        if (q) (void)hipFree(q);
    const Meas& m = c->meas;
    void* mptrs[] = {m.chunks, m.cinst, m.ichunk, m.part, m.coff, m.view, m.cross, m.k0, m.k1, m.v0, m.v1, m.khi,
                     m.scum, m.slot, m.mbase, m.moff, m.mp, m.mv, m.minst, m.midx, m.iparam, m.parent, m.label,
                     m.cflag, m.cstart, m.tcnt, m.cp, m.cv, m.crec, m.tpe, m.word, m.csum, m.sel, m.srec};
    for (void* q : mptrs)
        if (q) (void)hipFree(q);
    const Adj& a = c->adj;
    void* aptrs[] = {a.xb, a.vb, a.ub, a.mb, a.ab, a.sa, a.sxb, a.svb, a.srb, a.sqb, a.scb, a.nbar, a.pbar, a.red};
    for (void* q : aptrs)
        if (q) (void)hipFree(q);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->rb_event) (void)hipEventDestroy(c->rb_event);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

void alloc_all(isph_ctx* c, uint32_t cap) {
    Dev& d = c->d;
    d.cap = cap;
    size_t nodes = 2 * (size_t)cap;
    uint32_t sort_blocks = (cap + kSortTile - 1) / kSortTile;
    dalloc(&d.posh, cap);
    dalloc(&d.velu, cap);
    dalloc(&d.mass, cap);
    dalloc(&d.mat, cap);
    dalloc(&d.prov, cap);
    dalloc(&d.rho, cap);
    dalloc(&d.acc, cap);
    dalloc(&d.keys, cap);
    dalloc(&d.vals, cap);
    dalloc(&d.keys_alt, cap);
    dalloc(&d.vals_alt, cap);
    dalloc(&d.hist, 256 * ((size_t)sort_blocks + 1));  // counts per (digit, block), then the 256 digit totals
    d.nsort_blocks = sort_blocks;
    dalloc(&d.spos, cap);
    dalloc(&d.svel, cap);
    dalloc(&d.sh, cap);
    dalloc(&d.smat, cap);
    dalloc(&d.srho, cap);
    dalloc(&d.spr, cap);
    dalloc(&d.scs, cap);
    dalloc(&d.sgrav, cap);
    dalloc(&d.lcode, cap);
    dalloc(&d.left, cap);
    dalloc(&d.right, cap);
    dalloc(&d.parent, nodes);
    dalloc(&d.cslot, cap);
    dalloc(&d.nrange, cap);
    dalloc(&d.nlevel, cap + 4);
    d.carry_blocks = (cap + kUpLeaves - 1) / kUpLeaves;
    dalloc(&d.carry, 2 * (size_t)d.carry_blocks * kUpCarry);
    dalloc(&d.ncarry, 2 * (size_t)d.carry_blocks);
    dalloc(&d.groups, cap);
    dalloc(&d.ggroups, cap);
    {
        // expected ~0.25 chunks per particle, plus up to one part-used block per resident walk wave
        const size_t blocks = ((size_t)cap / 2 + kChunksPerBlock - 1) / kChunksPerBlock + kWalkWaves;
        dalloc(&d.chead, cap);
        dalloc(&d.cnext, blocks * kChunksPerBlock);
        dalloc(&d.ccnt, blocks * kChunksPerBlock);
        dalloc(&d.cpool, blocks * kChunksPerBlock * kWave);
        dalloc(&d.cmask, blocks * kChunksPerBlock * kWave);
        d.cblocks = (uint32_t)blocks;
    }
    dalloc(&d.gflag, cap);
    dalloc(&d.gscan, cap / kWave + 16);
    fill0(c->stream, d.gflag, (size_t)cap * sizeof(uint32_t));
    dalloc(&d.nG, 2 * nodes);

    dalloc(&d.nA, nodes);
    dalloc(&d.nS, nodes);
    dalloc(&d.nQ0, nodes);
    dalloc(&d.nQ1, nodes);
    dalloc(&d.scal, kScalars);
    dalloc(&d.dt, 1);
    dalloc(&d.time, 1);
    dalloc(&d.counters, (size_t)kCounterSlots * kWalkWaves);
    dalloc(&d.red, 1024 * 16);
    dalloc(&d.inject, inject_slots(d));
    fill0(c->stream, d.inject, inject_slots(d) * sizeof(double));
    dalloc(&d.eos, kMaxMaterials);
    dalloc(&d.ext, kMaxExtBodies);
    dalloc(&d.mtarget, cap);
    HIP_OK(hipMalloc(&d.aos, (size_t)cap * sizeof(isph_particle)));
    HIP_OK(hipHostMalloc((void**)&c->pinned, (size_t)cap * sizeof(isph_particle), hipHostMallocDefault));
    HIP_OK(hipEventCreateWithFlags(&c->rb_event, hipEventDisableTiming));
    fill0(c->stream, d.counters, (size_t)kCounterSlots * kWalkWaves * sizeof(unsigned long long));
    fill0(c->stream, d.time, sizeof(double));
    fill0(c->stream, d.dt, sizeof(float));
    fill0(c->stream, d.eos, kMaxMaterials * sizeof(Till));
    HIP_OK(hipStreamSynchronize(c->stream));
}

// pos64 marks the f64 state as present (the forms and getters test it), so it is allocated last.
void ensure64(isph_ctx* c) {
    if (!c->d.pos64) {
        dalloc_once(&c->d.vel64, 3 * (size_t)c->d.cap);
        dalloc_once(&c->d.srho64, c->d.cap);
        dalloc_once(&c->d.spr64, c->d.cap);
        dalloc_once(&c->d.scs64, c->d.cap);
        dalloc_once(&c->d.acc64, 3 * (size_t)c->d.cap);
        fill0(c->stream, c->d.acc64, 3 * (size_t)c->d.cap * sizeof(double));
        dalloc(&c->d.pos64, 3 * (size_t)c->d.cap);
    }
}

void ensure_ref(isph_ctx* c) {
    Dev& d = c->d;
    if (!d.slo) {
        dalloc_once(&d.svlo, d.cap);
        dalloc_once(&d.ndepth, d.cap);
        dalloc_once(&d.nA64, d.cap);
        dalloc_once(&d.sgrav64, d.cap);
        dalloc(&d.slo, d.cap);
    }
}

// ---- timing -------------------------------------------------------------------------------------------
hipEvent_t next_event(isph_ctx* c);

void flush_timing(isph_ctx* c) {
    if (c->ev_pairs.empty()) return;
    HIP_OK(hipStreamSynchronize(c->stream));
    for (const EvPair& e : c->ev_pairs) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e.a, e.b));
        c->phase_ms[e.phase] += ms;
    }
    c->ev_pairs.clear();
    c->ev_used = 0;
}

hipEvent_t next_event(isph_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        if (c->ev_pool.size() < 16384) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            c->ev_pool.push_back(e);
        } else {
            flush_timing(c);
        }
    }
    return c->ev_pool[c->ev_used++];
}

struct Phase {
    isph_ctx* c;
    int phase;
    hipEvent_t a = nullptr;
    Phase(isph_ctx* c_, int ph) : c(c_), phase(ph) {
        if (c->timing) {
            a = next_event(c);
            HIP_OK(hipEventRecord(a, c->stream));
        }
    }
    ~Phase() {
        if (!c->timing || !a) return;
        try {
            hipEvent_t b = next_event(c);
            if (hipEventRecord(b, c->stream) == hipSuccess) c->ev_pairs.push_back({phase, a, b});
        } catch (...) {
            // a failed timing record drops this phase sample; the work itself was enqueued
        }
    }
};

// ---- pipeline -----------------------------------------------------------------------------------------
void set_leaves(isph_ctx* c, uint32_t n) {
    c->d.n = n;
    c->d.L = (n + c->d.K - 1) / c->d.K;
    if (c->d.L == 0) c->d.L = 1;
}

void build_sorted(isph_ctx* c, const Dev& d, const float4* pos, const float* mass, bool with_sph) {
    Phase ph(c, ISPH_PHASE_SORT);
    launch_reset_scalars(d, c->stream);
    if (d.ninst > 1) {  // ensemble: per-instance boxes, (instance, code) order
        launch_reset_seg(d, c->stream, 0);
        launch_sorted_seg(d, pos, mass, c->stream);
        return;
    }
    launch_bbox(d, pos, c->stream);
    launch_morton(d, pos, c->stream);
    launch_radix_sort(d, c->stream);
    Dev g = d;
    if (!with_sph) {
        g.sh = nullptr;
        g.svel = nullptr;
        g.smat = nullptr;
    }
    launch_gather(g, pos, mass, c->stream);
}

void build_tree(isph_ctx* c, const Dev& d) {
    {
        Phase ph(c, ISPH_PHASE_TREE);
        launch_karras(d, c->stream);
        if (d.ninst > 1) launch_inst_roots(d, c->stream);
    }
    Phase ph(c, ISPH_PHASE_UPSWEEP);
    launch_upsweep(d, c->p, c->stream);
    launch_groups(d, c->p, kGroupMax, d.groups, d.scal + kGroupsSph, c->stream);
    if (c->p.grav_ppl == 2)
        launch_groups(d, c->p, 2 * kGroupMax, d.ggroups, d.scal + kGroupsGrav, c->stream);
}

void run_gravity(isph_ctx* c, const Dev& d, int mode, bool ref = false) {
    Phase ph(c, ISPH_PHASE_GRAVITY);
    if (ref && mode == ISPH_GRAVITY_TREE) {
        launch_com64(d, c->stream);
        launch_gravity_tree_ref(d, c->p, c->stream);
    } else if (ref && mode == ISPH_GRAVITY_DIRECT) {
        launch_gravity_direct_ref(d, c->p, c->stream);
    } else if (ref) {
        launch_gravity_zero_ref(d, c->stream);
    } else if (mode == ISPH_GRAVITY_TREE) {
        launch_gravity_tree(d, c->p, c->stream);
        c->grav_launches += 1;
    } else if (mode == ISPH_GRAVITY_DIRECT) {
        launch_gravity_direct(d, c->p, c->stream);
        c->grav_launches += 1;
    } else {
        launch_gravity_zero(d, c->stream);
    }
}

// One force evaluation (gpu_sph.rs:1468-1474): sort + tree + gravity + density/EOS + force.
void eval(isph_ctx* c, bool rebuild) {
    const Dev& d = c->d;
    const bool ref = c->p.ref64 != 0;
    if (ref && (d.K != 1 || c->p.quad || c->p.grav_ppl != 1 || d.ninst > 1 || !d.pos64))
        throw std::runtime_error("the f64 reference form takes one system, leaf size 1, monopole gravity and "
                                 "one particle per lane");
    if (rebuild) {
        build_sorted(c, d, d.posh, d.mass, true);
        if (ref) launch_gather_lo(d, c->stream);
        build_tree(c, d);
        run_gravity(c, d, c->gravity_mode, ref);
        Phase ph(c, ISPH_PHASE_DENSITY);
        launch_density(d, c->p, c->stream);
    } else {
        Phase ph(c, ISPH_PHASE_DENSITY);
        launch_reset_scalars(d, c->stream, 1);
        if (d.ninst > 1) launch_reset_seg(d, c->stream, 1);
        launch_gather_vel(d, c->stream);
        if (ref) launch_gather_lo(d, c->stream);
        launch_eos(d, c->stream);
    }
    Phase ph(c, ISPH_PHASE_FORCE);
    launch_force(d, c->p, c->stream);
}

void kdk(isph_ctx* c, bool cfl_mode, float cfl) {
    const Dev& d = c->d;
    eval(c, !(c->reuse_enabled && c->reuse_valid));
    {
        Phase ph(c, ISPH_PHASE_INTEGRATE);
        if (cfl_mode) {
            if (d.ninst > 1) launch_set_dt_cfl_seg(d, cfl, c->stream);
            else launch_set_dt_cfl(d, cfl, c->stream);
        }
        launch_kick_drift(d, c->p, c->stream);
    }
    eval(c, true);
    {
        Phase ph(c, ISPH_PHASE_INTEGRATE);
        launch_kick(d, c->stream);
        if (d.ninst > 1) launch_advance_time_seg(d, c->stream);
        else launch_advance_time(d, c->stream);
    }
    c->reuse_valid = true;
}

void drop_graph(isph_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_mode = -1;
}

// One KDK step, launched from the captured graph when graph mode is on (capturing it on first use).
// The step must be in its steady state when captured: with evaluation reuse on, the first step after a
// change runs eagerly so that the captured step is the reusing one.
void step_once(isph_ctx* c, bool cfl_mode, float value) {
    const Dev& d = c->d;
    auto body = [&]() {
        if (!cfl_mode) {
            if (d.ninst > 1) launch_set_dt_value_seg(d, value, c->stream);
            else launch_set_dt_value(d, value, c->stream);
        }
        kdk(c, cfl_mode, value);
    };
    if (!c->graph_on || (c->reuse_enabled && !c->reuse_valid)) {
        body();
        return;
    }
    const int mode = cfl_mode ? 1 : 0;
    if (!c->graph_exec || c->graph_mode != mode || c->graph_value != value) {
        drop_graph(c);
        const bool timing = c->timing;
        c->timing = false;  // no event records inside the graph
        HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            c->timing = timing;
            throw;
        }
        c->timing = timing;
        HIP_OK(hipStreamEndCapture(c->stream, &c->graph));
        HIP_OK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
        c->graph_mode = mode;
        c->graph_value = value;
    }
    Phase ph(c, ISPH_PHASE_INTEGRATE);  // with timing on, the whole replayed step is one sample
    HIP_OK(hipGraphLaunch(c->graph_exec, c->stream));
}

int fail(isph_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

template <class F>
int guarded(isph_ctx* c, F&& f) {
    if (!c) return ISPH_EINVAL;
    try {
        return f();
    } catch (const HipError& e) {
        return fail(c, ISPH_EDEVICE, e.what());
    } catch (const std::bad_alloc&) {
        return fail(c, ISPH_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(c, ISPH_EINVAL, e.what());
    }
}

void widen_eos(isph_ctx* c, const isph_eos* eos, size_t n) {
    Till t[kMaxMaterials];
    std::memset(t, 0, sizeof(t));
    for (size_t i = 0; i < n; ++i) {
        t[i] = Till{eos[i].rho0, eos[i].a, eos[i].b, eos[i].cap_a, eos[i].cap_b, eos[i].e0, eos[i].e_iv, eos[i].e_cv,
                    eos[i].alpha, eos[i].beta};
    }
    h2d(c->stream, c->d.eos, t, sizeof(t));
}

// ---- adjoint driver --------------------------------------------------------------------------------
void ensure_adj(isph_ctx* c) {
    Adj& a = c->adj;
    if (a.xb) return;
    const size_t cap = c->d.cap;
    dalloc_once(&a.vb, 3 * cap);
    dalloc_once(&a.ub, cap);
    dalloc_once(&a.mb, cap);
    dalloc_once(&a.ab, cap);
    dalloc_once(&a.sa, cap);
    dalloc_once(&a.sxb, cap);
    dalloc_once(&a.svb, cap);
    dalloc_once(&a.srb, cap);
    dalloc_once(&a.sqb, cap);
    dalloc_once(&a.scb, cap);
    dalloc_once(&a.nbar, kNbarSlots * 2 * (size_t)cap);
    dalloc_once(&a.pbar, kMaxMaterials * 4);
    dalloc_once(&a.red, 16);
    dalloc(&a.xb, 3 * cap);  // tested above: last
}

// The per-evaluation arrays of Dev: one set per stored evaluation of a checkpoint segment.
struct EvalBufs {
    uint32_t *keys = nullptr, *vals = nullptr, *keys_alt = nullptr, *vals_alt = nullptr, *hist = nullptr;
    float4 *spos = nullptr, *svel = nullptr, *sgrav = nullptr, *acc = nullptr;
    float *sh = nullptr, *srho = nullptr, *spr = nullptr, *scs = nullptr;
    uint32_t* smat = nullptr;
    double *srho64 = nullptr, *spr64 = nullptr, *scs64 = nullptr;
    uint32_t *lcode = nullptr, *left = nullptr, *right = nullptr, *parent = nullptr, *cslot = nullptr;
    uint8_t* nlevel = nullptr;
    uint32_t *gscan = nullptr, *scal = nullptr;
    uint2 *nrange = nullptr, *groups = nullptr;
    SphNode* nS = nullptr;
    uint4* nG = nullptr;
    float4* nA = nullptr;

    void alloc(uint32_t cap, uint32_t sort_blocks) {
        const size_t nodes = 2 * (size_t)cap;
        dalloc(&keys, cap); dalloc(&vals, cap); dalloc(&keys_alt, cap); dalloc(&vals_alt, cap);
        dalloc(&hist, 256 * ((size_t)sort_blocks + 1));
        dalloc(&spos, cap); dalloc(&svel, cap); dalloc(&sgrav, cap); dalloc(&acc, cap);
        dalloc(&sh, cap); dalloc(&srho, cap); dalloc(&spr, cap); dalloc(&scs, cap); dalloc(&smat, cap);
        dalloc(&srho64, cap); dalloc(&spr64, cap); dalloc(&scs64, cap);
        dalloc(&lcode, cap); dalloc(&left, cap); dalloc(&right, cap); dalloc(&parent, nodes); dalloc(&cslot, cap);
        dalloc(&nlevel, cap + 4); dalloc(&gscan, cap / kWave + 16); dalloc(&scal, kScalars);
        dalloc(&nrange, cap); dalloc(&groups, cap);
        dalloc(&nG, 2 * nodes); dalloc(&nA, nodes); dalloc(&nS, nodes);
    }
    void release() {
        void* ptrs[] = {keys, vals, keys_alt, vals_alt, hist, spos, svel, sgrav, acc, sh, srho, spr, scs, smat,
                        srho64, spr64, scs64, lcode, left, right, parent, cslot, nlevel, gscan, scal, nrange, groups,
                        nS, nG, nA};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        *this = EvalBufs{};
    }
    // Exchange these arrays with the ones the pipeline writes.
    void swap_with(Dev& d) {
        std::swap(keys, d.keys); std::swap(vals, d.vals); std::swap(keys_alt, d.keys_alt);
        std::swap(vals_alt, d.vals_alt); std::swap(hist, d.hist);
        std::swap(spos, d.spos); std::swap(svel, d.svel); std::swap(sgrav, d.sgrav); std::swap(acc, d.acc);
        std::swap(sh, d.sh); std::swap(srho, d.srho); std::swap(spr, d.spr); std::swap(scs, d.scs);
        std::swap(smat, d.smat); std::swap(srho64, d.srho64); std::swap(spr64, d.spr64); std::swap(scs64, d.scs64);
        std::swap(lcode, d.lcode); std::swap(left, d.left); std::swap(right, d.right); std::swap(parent, d.parent);
        std::swap(cslot, d.cslot); std::swap(nlevel, d.nlevel); std::swap(gscan, d.gscan); std::swap(scal, d.scal);
        std::swap(nrange, d.nrange); std::swap(groups, d.groups); std::swap(nS, d.nS); std::swap(nG, d.nG);
        std::swap(nA, d.nA);
    }
};

// Snapshot of the integrated state: f64 position and velocity, f32 internal energy.
struct StateBuf {
    double *x = nullptr, *v = nullptr;
    float* u = nullptr;
    void alloc(uint32_t cap) {
        dalloc(&x, 3 * (size_t)cap);
        dalloc(&v, 3 * (size_t)cap);
        dalloc(&u, cap);
    }
    void release() {
        if (x) (void)hipFree(x);
        if (v) (void)hipFree(v);
        if (u) (void)hipFree(u);
        *this = StateBuf{};
    }
};

// Owns every buffer of one reverse pass; frees them on any exit.
struct VjpScratch {
    std::vector<StateBuf> ck, start, mid;
    std::vector<EvalBufs> ev;
    float* dts = nullptr;
    StateBuf final_state;
    ~VjpScratch() {
        for (auto& b : ck) b.release();
        for (auto& b : start) b.release();
        for (auto& b : mid) b.release();
        for (auto& e : ev) e.release();
        final_state.release();
        if (dts) (void)hipFree(dts);
    }
};

// One force evaluation into the arrays of `e` (they are swapped in for the evaluation and back out).
void eval_into(isph_ctx* c, EvalBufs& e) {
    e.swap_with(c->d);
    try {
        eval(c, true);
    } catch (...) {
        e.swap_with(c->d);
        throw;
    }
    e.swap_with(c->d);
    // the integrators read the evaluation's output from the handle's own array
    HIP_OK(hipMemcpyAsync(c->d.acc, e.acc, (size_t)c->d.n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
}

void vjp_of(isph_ctx* c, EvalBufs& e) {
    e.swap_with(c->d);
    try {
        launch_vjp_eval(c->d, c->p, c->gravity_mode, c->adj, c->stream);
    } catch (...) {
        e.swap_with(c->d);
        throw;
    }
    e.swap_with(c->d);
}

}  // namespace

extern "C" {

isph_t* isph_create(uint32_t capacity, int device) {
    isph_ctx* c = nullptr;
    try {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            g_create_error = "no HIP device";
            return nullptr;
        }
        if (device < 0 || device >= count) {
            g_create_error = "device ordinal out of range";
            return nullptr;
        }
        c = new isph_ctx();
        c->device = device;
        HIP_OK(hipSetDevice(device));
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        alloc_all(c, std::max<uint32_t>(capacity, 1));
        c->d.K = 1;
        set_leaves(c, 0);
        g_create_error.clear();
        return c;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        if (c) {
            free_all(c);
            delete c;
        }
        return nullptr;
    }
}

void isph_destroy(isph_t* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
}

const char* isph_last_error(const isph_t* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

uint32_t isph_count(const isph_t* h) { return h ? h->d.n : 0; }

int isph_upload(isph_t* h, const isph_particle* particles, size_t n, const isph_eos* eos, size_t n_eos,
                float softening) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        if (n > 0 && !particles) return fail(h, ISPH_EINVAL, "particles is NULL");
        if (n_eos > 0 && !eos) return fail(h, ISPH_EINVAL, "eos is NULL");
        size_t ne = std::min<size_t>(n_eos, kMaxMaterials);  // gpu_sph.rs:1337
        size_t take = std::min<size_t>(n, h->d.cap);          // gpu_store.rs:33
        for (size_t i = 0; i < take; ++i)
            if (particles[i].mat >= ne) return fail(h, ISPH_EINVAL, "particle material index outside the EOS table");
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        if (take > 0) {
            h2d(h->stream, h->d.aos, particles, take * sizeof(isph_particle));
            launch_unpack(h->d, (uint32_t)take, h->stream);
        }
        if (ne > 0) widen_eos(h, eos, ne);
        h->n_eos = (uint32_t)ne;
        set_leaves(h, (uint32_t)take);
        h->d.ninst = 1;
        h->hioff = {0u, (uint32_t)take};
        h->meas.valid = false;
        h->softening = softening;
        h->p.soft2 = softening * softening;
        h->p.soft2_64 = (double)softening * (double)softening;
        if (h->precision != ISPH_PRECISION_FP32 && take > 0) {
            ensure64(h);
            launch_sync64_from32(h->d, h->stream);
        }
        h->reuse_valid = false;
        h->rb_inflight = false;
        HIP_OK(hipStreamSynchronize(h->stream));
        return (int)take;
    });
}

int isph_upload_ensemble(isph_t* h, const isph_particle* particles, const uint32_t* offsets, uint32_t n_inst,
                         const isph_eos* eos, size_t n_eos, float softening) {
    if (!h) return ISPH_EINVAL;
    if (!offsets || n_inst == 0 || n_inst > 256) return fail(h, ISPH_EINVAL, "ensemble of 1..256 instances expected");
    if (offsets[0] != 0) return fail(h, ISPH_EINVAL, "offsets[0] must be 0");
    for (uint32_t k = 0; k < n_inst; ++k)
        if (offsets[k + 1] <= offsets[k]) return fail(h, ISPH_EINVAL, "every instance needs at least one particle");
    const size_t n = offsets[n_inst];
    if (n > h->d.cap) return fail(h, ISPH_ERANGE, "ensemble exceeds the handle capacity");
    if (h->d.K != 1) return fail(h, ISPH_EINVAL, "ensembles use leaf size 1");
    if (n_inst > 1 && h->gravity_mode == ISPH_GRAVITY_DIRECT)
        return fail(h, ISPH_EINVAL, "ensembles use tree (or no) gravity: the direct sum would couple instances");
    int rc = isph_upload(h, particles, n, eos, n_eos, softening);
    if (rc < 0) return rc;
    return guarded(h, [&]() -> int {
        Dev& d = h->d;
        void* old[] = {d.inst, d.mcode, d.lkey, d.ioff, d.ibox, d.iroot, d.isig, d.idt, d.itime};
        for (void* q : old)
            if (q) (void)hipFree(q);
        d.inst = d.mcode = d.ioff = d.ibox = d.iroot = d.isig = nullptr;
        d.lkey = nullptr;
        d.idt = nullptr;
        d.itime = nullptr;
        d.ninst = 1;
        drop_graph(h);
        if (n_inst == 1) return (int)n;
        dalloc(&d.inst, d.cap);
        dalloc(&d.mcode, d.cap);
        dalloc(&d.lkey, d.cap);
        dalloc(&d.ioff, n_inst + 1);
        dalloc(&d.ibox, 6 * (size_t)n_inst);
        dalloc(&d.iroot, n_inst);
        dalloc(&d.isig, n_inst);
        dalloc(&d.idt, n_inst);
        dalloc(&d.itime, n_inst);
        std::vector<uint32_t> inst(n);
        for (uint32_t k = 0; k < n_inst; ++k)
            for (uint32_t i = offsets[k]; i < offsets[k + 1]; ++i) inst[i] = k;
        h2d(h->stream, d.inst, inst.data(), n * sizeof(uint32_t));
        h2d(h->stream, d.ioff, offsets, (n_inst + 1) * sizeof(uint32_t));
        fill0(h->stream, d.idt, n_inst * sizeof(float));
        fill0(h->stream, d.itime, n_inst * sizeof(double));
        d.ninst = n_inst;
        h->hioff.assign(offsets, offsets + n_inst + 1);
        h->reuse_valid = false;
        drop_graph(h);
        return (int)n;
    });
}

int isph_ensemble_times(isph_t* h, double* t, float* dt, uint32_t cap) {
    return guarded(h, [&]() -> int {
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        if (h->d.ninst <= 1) {
            if (cap >= 1 && t) HIP_OK(hipMemcpy(t, h->d.time, sizeof(double), hipMemcpyDeviceToHost));
            if (cap >= 1 && dt) HIP_OK(hipMemcpy(dt, h->d.dt, sizeof(float), hipMemcpyDeviceToHost));
            return 1;
        }
        const size_t k = std::min<size_t>(cap, h->d.ninst);
        if (t && k) HIP_OK(hipMemcpy(t, h->d.itime, k * sizeof(double), hipMemcpyDeviceToHost));
        if (dt && k) HIP_OK(hipMemcpy(dt, h->d.idt, k * sizeof(float), hipMemcpyDeviceToHost));
        return (int)h->d.ninst;
    });
}

int isph_set_eos_f64(isph_t* h, const double* params10, size_t n_eos) {
    return guarded(h, [&]() -> int {
        if (!params10 || n_eos == 0 || n_eos > (size_t)kMaxMaterials) return fail(h, ISPH_EINVAL, "bad EOS table");
        Till t[kMaxMaterials];
        std::memset(t, 0, sizeof(t));
        for (size_t i = 0; i < n_eos; ++i) {
            const double* q = params10 + 10 * i;
            t[i] = Till{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9]};
        }
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        h2d(h->stream, h->d.eos, t, sizeof(t));
        h->n_eos = (uint32_t)n_eos;
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_set_dt(isph_t* h, float dt, float damp) {
    return guarded(h, [&]() -> int {
        h->dt_set = dt;
        h->damp = damp;
        return ISPH_OK;
    });
}

int isph_set_av(isph_t* h, float alpha, float beta) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        h->p.av_alpha = alpha;
        h->p.av_beta = beta;
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_set_gravity(isph_t* h, int mode, float theta, uint32_t leaf_size, int quadrupole) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        if (mode < ISPH_GRAVITY_TREE || mode > ISPH_GRAVITY_NONE) return fail(h, ISPH_EINVAL, "bad gravity mode");
        if (!(theta > 0.0f) || leaf_size < 1 || leaf_size > 64) return fail(h, ISPH_EINVAL, "bad theta or leaf size");
        if (h->d.ninst > 1 && (leaf_size != 1 || mode == ISPH_GRAVITY_DIRECT))
            return fail(h, ISPH_EINVAL, "ensembles use leaf size 1 and tree (or no) gravity");
        h->gravity_mode = mode;
        h->p.theta = theta;
        h->p.quad = quadrupole ? 1 : 0;
        h->d.K = leaf_size;
        set_leaves(h, h->d.n);
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_set_walk(isph_t* h, float group_extent, int grav_particles_per_lane, int sph_expand) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        if (!(group_extent > 0.0f) || (grav_particles_per_lane != 1 && grav_particles_per_lane != 2) ||
            sph_expand < 1 || sph_expand > 64)
            return fail(h, ISPH_EINVAL, "bad walk shape");
        h->p.group_extent = group_extent;
        h->p.grav_ppl = grav_particles_per_lane;
        h->p.sph_expand = sph_expand;
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_set_candidate_cache(isph_t* h, uint32_t max_blocks) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        h->p.cache_blocks = max_blocks;
        return ISPH_OK;
    });
}

int isph_set_precision(isph_t* h, int mode) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        if (mode != ISPH_PRECISION_FP32 && mode != ISPH_PRECISION_MIXED && mode != ISPH_PRECISION_F64)
            return fail(h, ISPH_EINVAL, "bad precision");
        HIP_OK(hipSetDevice(h->device));
        if (mode != ISPH_PRECISION_FP32) {
            ensure64(h);
            if (mode == ISPH_PRECISION_F64) ensure_ref(h);
            // entering an f64 form from the plain one takes the f32 state; between f64 forms the state stays
            if (h->d.n > 0 && h->precision == ISPH_PRECISION_FP32) launch_sync64_from32(h->d, h->stream);
            h->p.mixed = 1;
        } else {
            h->p.mixed = 0;
        }
        h->p.ref64 = mode == ISPH_PRECISION_F64 ? 1 : 0;
        h->precision = mode;
        // The integrators take the f64 arrays only in mixed mode.
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

static Dev integ_view(isph_ctx* c);

int isph_set_bulk(isph_t* h, double cx, double cy, double cz, double r_core, double vx, double vy, double vz,
                  double mass) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        h->p.bulk_cr = make_float4((float)cx, (float)cy, (float)cz, (float)r_core);  // gpu_sph.rs:1372-1378
        h->p.bulk_vm = make_float4((float)vx, (float)vy, (float)vz, (float)mass);
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_set_bulk_planar(isph_t* h, double px, double py, double pz, double ux, double uy, double uz, double g) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        double l = std::sqrt(ux * ux + uy * uy + uz * uz);
        if (!(l > 0.0) || !std::isfinite(l)) {  // normalize_or(+Y), gpu_sph.rs:1394
            ux = 0.0;
            uy = 1.0;
            uz = 0.0;
            l = 1.0;
        }
        h->p.bulk_cr = make_float4((float)px, (float)py, (float)pz, -1.0f);
        h->p.bulk_vm = make_float4((float)(ux / l), (float)(uy / l), (float)(uz / l), (float)g);
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_set_external_masses(isph_t* h, const double* pos_mass4, size_t n) {
    return guarded(h, [&]() -> int {
        if (n > 0 && !pos_mass4) return fail(h, ISPH_EINVAL, "NULL bodies");
        const size_t k = std::min<size_t>(n, kMaxExtBodies);  // overflow dropped, count returned (:1417)
        std::vector<float4> packed(k);
        for (size_t i = 0; i < k; ++i)
            packed[i] = make_float4((float)pos_mass4[4 * i], (float)pos_mass4[4 * i + 1], (float)pos_mass4[4 * i + 2],
                                    (float)pos_mass4[4 * i + 3]);
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        if (k) h2d(h->stream, h->d.ext, packed.data(), k * sizeof(float4));
        drop_graph(h);
        h->p.n_ext = (uint32_t)k;
        h->reuse_valid = false;
        return (int)k;
    });
}

int isph_set_merge_budget(isph_t* h, uint32_t budget) {
    return guarded(h, [&]() -> int {
        h->p.merge_budget = budget;
        return ISPH_OK;
    });
}

int isph_set_omega(isph_t* h, float omega) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        h->p.omega = omega;
        return ISPH_OK;
    });
}

int isph_encode_relax(isph_t* h, uint32_t steps) {
    return guarded(h, [&]() -> int {
        if (h->d.n == 0 || steps == 0) return ISPH_OK;
        if (h->d.ninst > 1) return fail(h, ISPH_EINVAL, "relaxation takes one system");
        HIP_OK(hipSetDevice(h->device));
        Dev saved = h->d;
        h->d = integ_view(h);
        try {
            launch_set_dt_value(h->d, h->dt_set, h->stream);
            for (uint32_t s = 0; s < steps; ++s) {
                eval(h, true);  // one force evaluation (gpu_sph.rs:1476-1481)
                Phase ph(h, ISPH_PHASE_INTEGRATE);
                launch_relax(h->d, h->p, h->damp, h->stream);
            }
        } catch (...) {
            h->d = saved;
            throw;
        }
        h->d = saved;
        h->reuse_valid = false;
        HIP_OK(hipGetLastError());
        return ISPH_OK;
    });
}

int isph_encode_merge(isph_t* h) {
    return guarded(h, [&]() -> int {
        // gpu_sph.rs:1439-1447: neighbours + density, then pick -> apply -> retire; a no-op unless the
        // particle count exceeds the budget (sph_step.wgsl:386-388)
        if (h->d.n == 0 || h->p.merge_budget == 0 || h->d.n <= h->p.merge_budget) return ISPH_OK;
        HIP_OK(hipSetDevice(h->device));
        Dev saved = h->d;
        h->d = integ_view(h);
        try {
            build_sorted(h, h->d, h->d.posh, h->d.mass, true);
            build_tree(h, h->d);
            Phase ph(h, ISPH_PHASE_DENSITY);
            StepParams pm = h->p;
            pm.ref64 = 0;  // the merge neighbours and density need no f64 coordinates
            launch_density(h->d, pm, h->stream);
            launch_merge(h->d, pm, h->stream);
        } catch (...) {
            h->d = saved;
            throw;
        }
        h->d = saved;
        h->reuse_valid = false;
        HIP_OK(hipGetLastError());
        return ISPH_OK;
    });
}

int isph_set_graph(isph_t* h, int enable) {
    return guarded(h, [&]() -> int {
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        drop_graph(h);
        h->graph_on = enable != 0;
        return ISPH_OK;
    });
}

int isph_set_reuse(isph_t* h, int enable) {
    return guarded(h, [&]() -> int {
        drop_graph(h);
        h->reuse_enabled = enable != 0;
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

static Dev integ_view(isph_ctx* c) {
    Dev d = c->d;
    if (c->precision == ISPH_PRECISION_FP32) {
        d.pos64 = nullptr;
        d.vel64 = nullptr;
        d.srho64 = nullptr;
        d.spr64 = nullptr;
        d.scs64 = nullptr;
        d.acc64 = nullptr;
    }
    return d;
}

int isph_encode_kdk(isph_t* h, uint32_t substeps) {
    return guarded(h, [&]() -> int {
        if (h->d.n == 0 || substeps == 0) return ISPH_OK;
        HIP_OK(hipSetDevice(h->device));
        Dev saved = h->d;
        h->d = integ_view(h);
        try {
            for (uint32_t s = 0; s < substeps; ++s) step_once(h, false, h->dt_set);
        } catch (...) {
            h->d = saved;
            throw;
        }
        h->d = saved;
        HIP_OK(hipGetLastError());
        return ISPH_OK;
    });
}

int isph_step_cfl(isph_t* h, float cfl, uint32_t steps) {
    return guarded(h, [&]() -> int {
        if (!(cfl > 0.0f)) return fail(h, ISPH_EINVAL, "cfl must be > 0");
        if (h->d.n == 0 || steps == 0) return ISPH_OK;
        HIP_OK(hipSetDevice(h->device));
        Dev saved = h->d;
        h->d = integ_view(h);
        try {
            for (uint32_t s = 0; s < steps; ++s) step_once(h, true, cfl);
        } catch (...) {
            h->d = saved;
            throw;
        }
        h->d = saved;
        HIP_OK(hipGetLastError());
        return ISPH_OK;
    });
}

int isph_sim_time(isph_t* h, double* t, float* last_dt) {
    return guarded(h, [&]() -> int {
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        if (t) HIP_OK(hipMemcpy(t, h->d.time, sizeof(double), hipMemcpyDeviceToHost));
        if (last_dt) HIP_OK(hipMemcpy(last_dt, h->d.dt, sizeof(float), hipMemcpyDeviceToHost));
        return ISPH_OK;
    });
}

int isph_begin_readback(isph_t* h) {
    return guarded(h, [&]() -> int {
        if (h->d.n == 0 || h->rb_inflight) return ISPH_OK;  // gpu_store.rs:127-129
        HIP_OK(hipSetDevice(h->device));
        launch_pack(h->d, h->stream);
        HIP_OK(hipMemcpyAsync(h->pinned, h->d.aos, (size_t)h->d.n * sizeof(isph_particle), hipMemcpyDeviceToHost,
                              h->stream));
        HIP_OK(hipEventRecord(h->rb_event, h->stream));
        h->rb_inflight = true;
        h->rb_count = h->d.n;
        return ISPH_OK;
    });
}

int isph_take_readback(isph_t* h, isph_particle* out, size_t cap) {
    return guarded(h, [&]() -> int {
        if (!h->rb_inflight) return ISPH_EAGAIN;
        hipError_t q = hipEventQuery(h->rb_event);
        if (q == hipErrorNotReady) return ISPH_EAGAIN;
        HIP_OK(q);
        size_t m = std::min<size_t>(h->rb_count, cap);
        if (m > 0 && out) std::memcpy(out, h->pinned, m * sizeof(isph_particle));
        h->rb_inflight = false;
        return (int)h->rb_count;
    });
}

int isph_readback_sync(isph_t* h, isph_particle* out, size_t cap) {
    return guarded(h, [&]() -> int {
        HIP_OK(hipSetDevice(h->device));
        if (h->d.n == 0) return 0;
        launch_pack(h->d, h->stream);
        HIP_OK(hipMemcpyAsync(h->pinned, h->d.aos, (size_t)h->d.n * sizeof(isph_particle), hipMemcpyDeviceToHost,
                              h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        size_t m = std::min<size_t>(h->d.n, cap);
        if (m > 0 && out) std::memcpy(out, h->pinned, m * sizeof(isph_particle));
        return (int)h->d.n;
    });
}

int isph_particle_buffer(isph_t* h, void** dev_ptr, uint32_t* count) {
    return guarded(h, [&]() -> int {
        if (!dev_ptr) return fail(h, ISPH_EINVAL, "NULL output");
        HIP_OK(hipSetDevice(h->device));
        if (h->d.n > 0) launch_pack(h->d, h->stream);
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        *dev_ptr = h->d.aos;
        if (count) *count = h->d.n;
        return ISPH_OK;
    });
}

void* isph_stream(isph_t* h) { return h ? (void*)h->stream : nullptr; }

int isph_synchronize(isph_t* h) {
    return guarded(h, [&]() -> int {
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        return ISPH_OK;
    });
}

int isph_force_eval(isph_t* h, float* acc4, float* rho) {
    return guarded(h, [&]() -> int {
        if (h->d.n == 0) return ISPH_OK;
        HIP_OK(hipSetDevice(h->device));
        // The plain form's EOS pass must not take the f64 density a mixed evaluation left behind.
        Dev saved = h->d;
        h->d = integ_view(h);
        try {
            eval(h, true);
        } catch (...) {
            h->d = saved;
            throw;
        }
        h->d = saved;
        h->reuse_valid = false;
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        if (acc4) HIP_OK(hipMemcpy(acc4, h->d.acc, (size_t)h->d.n * sizeof(float4), hipMemcpyDeviceToHost));
        if (rho) HIP_OK(hipMemcpy(rho, h->d.rho, (size_t)h->d.n * sizeof(float), hipMemcpyDeviceToHost));
        return ISPH_OK;
    });
}

// ---- measurement -------------------------------------------------------------------------------------
namespace {

void ensure_meas(isph_ctx* h) {
    Meas& m = h->meas;
    if (m.cap == h->d.cap && m.word) return;
    const size_t cap = std::max<uint32_t>(h->d.cap, 1u);
    const size_t cap_inst = 256, cap_chunks = (cap + kMeasChunk - 1) / kMeasChunk + cap_inst;
    dalloc_once(&m.chunks, cap_chunks);
    dalloc_once(&m.cinst, cap_chunks);
    dalloc_once(&m.ichunk, cap_inst + 1);
    dalloc_once(&m.part, cap_chunks * kPartFields);
    dalloc_once(&m.coff, cap_chunks);
    dalloc_once(&m.view, cap_inst * kViewFields);
    dalloc_once(&m.cross, cap_inst);
    dalloc_once(&m.k0, cap);
    dalloc_once(&m.k1, cap);
    dalloc_once(&m.v0, cap);
    dalloc_once(&m.v1, cap);
    dalloc_once(&m.khi, cap);
    dalloc_once(&m.scum, cap);
    dalloc_once(&m.slot, cap);
    dalloc_once(&m.mbase, cap_chunks);
    dalloc_once(&m.moff, cap_inst + 1);
    dalloc_once(&m.mp, cap);
    dalloc_once(&m.mv, cap);
    dalloc_once(&m.minst, cap);
    dalloc_once(&m.midx, cap);
    dalloc_once(&m.iparam, cap_inst * kParamFields);
    dalloc_once(&m.parent, cap);
    dalloc_once(&m.label, cap);
    dalloc_once(&m.cflag, cap + 1);
    dalloc_once(&m.cstart, cap + 1);
    dalloc_once(&m.tcnt, cap + 1);
    dalloc_once(&m.cp, cap);
    dalloc_once(&m.cv, cap);
    dalloc_once(&m.crec, cap * kClumpFields);
    dalloc_once(&m.word, 24);
    dalloc_once(&m.csum, 3 * cap_inst);
    dalloc_once(&m.sel, cap);
    dalloc_once(&m.srec, cap * kClumpFields);
    m.cap = h->d.cap;
    m.cap_inst = (uint32_t)cap_inst;
    m.cap_chunks = (uint32_t)cap_chunks;
}

// Chunks of at most kMeasChunk particles that never span two instances.
void meas_chunks(isph_ctx* h, const std::vector<uint32_t>& ioff) {
    Meas& m = h->meas;
    const uint32_t ninst = (uint32_t)ioff.size() - 1;
    std::vector<uint2> ch;
    std::vector<uint32_t> ci, ic(ninst + 1);
    for (uint32_t k = 0; k < ninst; ++k) {
        ic[k] = (uint32_t)ch.size();
        for (uint32_t s = ioff[k]; s < ioff[k + 1]; s += kMeasChunk) {
            ch.push_back(make_uint2(s, std::min(s + kMeasChunk, ioff[k + 1])));
            ci.push_back(k);
        }
    }
    ic[ninst] = (uint32_t)ch.size();
    if (ch.size() > m.cap_chunks) throw std::runtime_error("measurement chunk table overflow");
    m.nchunks = (uint32_t)ch.size();
    if (!ch.empty()) {
        h2d(h->stream, m.chunks, ch.data(), ch.size() * sizeof(uint2));
        h2d(h->stream, m.cinst, ci.data(), ci.size() * sizeof(uint32_t));
    }
    h2d(h->stream, m.ichunk, ic.data(), ic.size() * sizeof(uint32_t));
}

uint32_t read_word(isph_ctx* h, int k) {
    uint32_t w = 0;
    HIP_OK(hipStreamSynchronize(h->stream));
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(&w, h->meas.word + k, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return w;
}

// Friends-of-friends, statistics, self energy and gates over members [0, n) of h->meas.
void run_clumps(isph_ctx* h, uint32_t n, uint32_t ninst) {
    Meas& m = h->meas;
    m.n_clumps = 0;
    if (n > 0) {
        launch_fof(m, n, h->d.hist, h->stream);
        if (read_word(h, 9) != 0) throw std::runtime_error("friends-of-friends: a union did not settle");
        const uint32_t nc = read_word(h, 8);
        launch_clump_stats(m, nc, h->stream);
        uint32_t tiles = 0;
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(&tiles, m.tcnt + nc, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (tiles > m.tpe_cap) {
            if (m.tpe) HIP_OK(hipFree(m.tpe));
            m.tpe = nullptr;
            m.tpe_cap = std::max<size_t>(tiles, 2 * m.tpe_cap);
            dalloc(&m.tpe, m.tpe_cap);
        }
        launch_clump_pe(m, nc, tiles, h->stream);
        m.n_clumps = nc;
    }
    launch_clump_summary(m, m.n_clumps, ninst, h->stream);
    HIP_OK(hipStreamSynchronize(h->stream));
    HIP_OK(hipGetLastError());
}

// The records of instance `inst`'s clumps with >= min_members members, heaviest first (ties by id), with
// ids made relative to `base`.
uint32_t fetch_clumps(isph_ctx* h, uint32_t inst, uint32_t min_members, uint32_t base, isph_clump* out,
                      uint32_t cap) {
    Meas& m = h->meas;
    launch_clump_select(m, m.n_clumps, inst, min_members, h->stream);
    const uint32_t count = read_word(h, 16);
    if (count == 0) return 0;
    launch_clump_gather(m, count, h->stream);
    std::vector<double> rec((size_t)count * kClumpFields);
    HIP_OK(hipStreamSynchronize(h->stream));
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(rec.data(), m.srec, rec.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<uint32_t> order(count);
    for (uint32_t k = 0; k < count; ++k) order[k] = k;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const double ma = rec[(size_t)a * kClumpFields + kCMass], mb = rec[(size_t)b * kClumpFields + kCMass];
        if (ma != mb) return ma > mb;
        return rec[(size_t)a * kClumpFields + kCId] < rec[(size_t)b * kClumpFields + kCId];
    });
    for (uint32_t k = 0; k < std::min(count, cap); ++k) {
        const double* r = &rec[(size_t)order[k] * kClumpFields];
        isph_clump& c = out[k];
        c.id = (uint32_t)r[kCId] - base;
        c.n_members = (uint32_t)r[kCN];
        c.bound = r[kCBound] != 0.0;
        c.outside_roche = r[kCOutside] != 0.0;
        c.mass = r[kCMass];
        for (int a = 0; a < 3; ++a) {
            c.com_pos[a] = r[kCCom + a];
            c.com_vel[a] = r[kCVel + a];
            c.ang_mom[a] = r[kCL + a];
        }
        c.rho = r[kCRho];
        c.radius = r[kCRadius];
        c.internal_ke = r[kCKe];
        c.self_pe = r[kCPe];
        c.thermal_ke = r[kCThermal];
    }
    return count;
}

}  // namespace

int isph_set_measure(isph_t* h, double clump_softening, double link_factor) {
    return guarded(h, [&]() -> int {
        if (!(clump_softening >= 0.0) || !(link_factor > 0.0)) return fail(h, ISPH_EINVAL, "bad measure parameters");
        h->clump_soft = clump_softening;
        h->link_factor = link_factor;
        return ISPH_OK;
    });
}

int isph_disk_view(isph_t* h, isph_disk* out, uint32_t cap) {
    return guarded(h, [&]() -> int {
        if (!out && cap > 0) return fail(h, ISPH_EINVAL, "NULL output");
        HIP_OK(hipSetDevice(h->device));
        Dev& d = h->d;
        Meas& m = h->meas;
        const uint32_t ninst = d.n == 0 ? 1u : d.ninst;
        m.valid = false;
        h->mview.assign((size_t)ninst * kViewFields, 0.0);
        h->mmoff.assign(ninst + 1, 0u);
        h->mcsum.assign(3 * (size_t)ninst, 0ull);
        if (d.n > 0) {
            ensure_meas(h);
            HIP_OK(hipStreamSynchronize(h->stream));
            meas_chunks(h, h->hioff);
            launch_meas_live(d, m, ninst, h->stream);
            launch_meas_remnant(d, m, ninst, h->stream);
            launch_meas_classify(d, m, ninst, h->clump_soft, h->link_factor, h->stream);
            m.n_members = read_word(h, 0);
            run_clumps(h, m.n_members, ninst);
            HIP_OK(hipMemcpy(h->mview.data(), m.view, h->mview.size() * sizeof(double), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(h->mmoff.data(), m.moff, h->mmoff.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(h->mcsum.data(), m.csum, h->mcsum.size() * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
        }
        m.ninst = ninst;
        m.valid = true;
        for (uint32_t k = 0; k < std::min(cap, ninst); ++k) {
            const double* v = &h->mview[(size_t)k * kViewFields];
            isph_disk& o = out[k];
            std::memset(&o, 0, sizeof(o));
            o.valid = v[kVValid] != 0.0;
            o.n_live = (uint32_t)v[kVNLive];
            if (!o.valid) continue;
            o.n_disk = (uint32_t)v[kVNDisk];
            const bool clumps = o.n_disk >= 2;  // gpu_sph.rs:794
            o.n_clumps = clumps ? (uint32_t)h->mcsum[3 * k] : 0u;
            o.n_bound = clumps ? (uint32_t)h->mcsum[3 * k + 1] : 0u;
            if (clumps) {
                const unsigned long long bits = h->mcsum[3 * k + 2];
                std::memcpy(&o.moon_mass, &bits, sizeof(double));
            }
            for (int a = 0; a < 3; ++a) {
                o.com[a] = v[kVCom + a];
                o.v_com[a] = v[kVVCom + a];
            }
            o.r_remnant = v[kVRRem];
            o.m_remnant = v[kVMRem];
            o.mu = v[kVMu];
            o.disk_prov0 = v[kVDisk0];
            o.disk_other = v[kVDiskO];
            o.escaped = v[kVEsc];
            o.mean_h = v[kVMeanH];
            o.remnant_prov0 = v[kVRem0];
            o.remnant_other = v[kVRemO];
            o.escaped_prov0 = v[kVEsc0];
        }
        return (int)ninst;
    });
}

namespace {
int disk_instance(isph_ctx* h, uint32_t inst) {
    if (!h->meas.valid) return fail(h, ISPH_EINVAL, "no disk view: call isph_disk_view first");
    if (inst >= h->meas.ninst) return fail(h, ISPH_ERANGE, "instance out of range");
    return ISPH_OK;
}
}  // namespace

int isph_disk_members(isph_t* h, uint32_t inst, uint32_t* idx, uint32_t* clump_id, uint32_t cap) {
    return guarded(h, [&]() -> int {
        int rc = disk_instance(h, inst);
        if (rc < 0) return rc;
        if (h->mview[(size_t)inst * kViewFields + kVValid] == 0.0) return 0;
        const uint32_t b = h->mmoff[inst], n = h->mmoff[inst + 1] - b, k = std::min(n, cap);
        HIP_OK(hipStreamSynchronize(h->stream));
        if (k > 0 && idx)
            HIP_OK(hipMemcpy(idx, h->meas.midx + b, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (k > 0 && clump_id) {
            HIP_OK(hipMemcpy(clump_id, h->meas.label + b, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < k; ++i) clump_id[i] -= b;
        }
        return (int)n;
    });
}

int isph_disk_clumps(isph_t* h, uint32_t inst, uint32_t min_members, isph_clump* out, uint32_t cap) {
    return guarded(h, [&]() -> int {
        int rc = disk_instance(h, inst);
        if (rc < 0) return rc;
        if (!out && cap > 0) return fail(h, ISPH_EINVAL, "NULL output");
        const double* v = &h->mview[(size_t)inst * kViewFields];
        if (v[kVValid] == 0.0 || v[kVNDisk] < 2.0) return 0;  // gpu_sph.rs:794
        HIP_OK(hipSetDevice(h->device));
        return (int)fetch_clumps(h, inst, std::max(min_members, 1u), h->mmoff[inst], out, cap);
    });
}

int isph_disk_stats_json(isph_t* h, uint32_t inst, char* buf, size_t cap) {
    return guarded(h, [&]() -> int {
        int rc = disk_instance(h, inst);
        if (rc < 0) return rc;
        const double* v = &h->mview[(size_t)inst * kViewFields];
        std::string txt = "null";
        if (v[kVValid] != 0.0) {  // gpu_sph.rs:825-851
            constexpr double kMMoon = 7.342e22;
            const double disk = v[kVDisk0] + v[kVDiskO];
            const double earth_pct = disk > 0.0 ? 100.0 * v[kVDisk0] / disk : 0.0;
            double moon = 0.0;
            if (v[kVNDisk] >= 2.0) std::memcpy(&moon, &h->mcsum[3 * (size_t)inst + 2], sizeof(double));
            char tmp[256];
            std::snprintf(tmp, sizeof(tmp),
                          "{\"disk\":%.3f,\"earth_pct\":%.0f,\"remnant_km\":%.0f,\"escaped\":%.3f,\"moon\":%.3f}",
                          disk / kMMoon, earth_pct, v[kVRRem] / 1e3, v[kVEsc] / kMMoon, moon / kMMoon);
            txt = tmp;
        }
        if (buf && cap > 0) {
            const size_t k = std::min(cap - 1, txt.size());
            std::memcpy(buf, txt.data(), k);
            buf[k] = '\0';
        }
        return (int)txt.size();
    });
}

int isph_find_clumps(isph_t* h, const double* pos3, const double* vel3, const double* mass, const double* rho,
                     size_t n, double linking_length, double g, double softening, const double* central_pos3,
                     double central_mass, double central_radius, uint32_t min_members, isph_clump* out,
                     uint32_t cap, uint32_t* labels) {
    return guarded(h, [&]() -> int {
        if (n > 0 && (!pos3 || !vel3 || !mass || !rho)) return fail(h, ISPH_EINVAL, "NULL array");
        if (!central_pos3) return fail(h, ISPH_EINVAL, "NULL central position");
        if (!out && cap > 0) return fail(h, ISPH_EINVAL, "NULL output");
        if (n > h->d.cap) return fail(h, ISPH_ERANGE, "n exceeds the handle capacity");
        if (!(linking_length >= 0.0) || !(softening >= 0.0)) return fail(h, ISPH_EINVAL, "bad linking length");
        HIP_OK(hipSetDevice(h->device));
        ensure_meas(h);
        Meas& m = h->meas;
        m.valid = false;
        HIP_OK(hipStreamSynchronize(h->stream));
        if (n == 0) return 0;
        std::vector<double4> p(n), v(n);
        for (size_t i = 0; i < n; ++i) {
            p[i] = make_double4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], mass[i]);
            v[i] = make_double4(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2], rho[i]);
        }
        double pr[kParamFields] = {0};
        pr[kPLink] = linking_length;
        pr[kPSoft2] = softening * softening;
        pr[kPG] = g;
        for (int a = 0; a < 3; ++a) pr[kPCPos + a] = central_pos3[a];
        pr[kPCMass] = central_mass;
        pr[kPCRad] = central_radius;
        h2d(h->stream, m.mp, p.data(), n * sizeof(double4));
        h2d(h->stream, m.mv, v.data(), n * sizeof(double4));
        fill0(h->stream, m.minst, n * sizeof(uint32_t));
        h2d(h->stream, m.iparam, pr, sizeof(pr));
        run_clumps(h, (uint32_t)n, 1u);
        HIP_OK(hipStreamSynchronize(h->stream));
        if (labels) HIP_OK(hipMemcpy(labels, m.label, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        return (int)fetch_clumps(h, 0u, std::max(min_members, 1u), 0u, out, cap);
    });
}

int isph_gravity_accelerations(isph_t* h, const double* pos3, const double* mass, size_t n, double softening,
                               double* acc3, int mode) {
    return guarded(h, [&]() -> int {
        if (n == 0) return ISPH_OK;
        if (!pos3 || !mass || !acc3) return fail(h, ISPH_EINVAL, "NULL array");
        if (n > h->d.cap) return fail(h, ISPH_ERANGE, "n exceeds the handle capacity");
        if (mode != ISPH_GRAVITY_TREE && mode != ISPH_GRAVITY_DIRECT) return fail(h, ISPH_EINVAL, "bad mode");
        HIP_OK(hipSetDevice(h->device));
        if (!h->gpos) {
            dalloc_once(&h->gmass, h->d.cap);
            dalloc(&h->gpos, h->d.cap);
        }
        // gpu_gravity.rs:340-345: bodies as f32 (pos.x, pos.y, pos.z, mass); soft2 = (s*s) as f32.
        std::vector<float4> b(n);
        std::vector<float> m(n);
        for (size_t i = 0; i < n; ++i) {
            b[i] = make_float4((float)pos3[3 * i], (float)pos3[3 * i + 1], (float)pos3[3 * i + 2], 0.f);
            m[i] = (float)mass[i];
        }
        HIP_OK(hipStreamSynchronize(h->stream));
        h2d(h->stream, h->gpos, b.data(), n * sizeof(float4));
        h2d(h->stream, h->gmass, m.data(), n * sizeof(float));
        Dev g = h->d;
        g.n = (uint32_t)n;
        g.L = (g.n + g.K - 1) / g.K;
        g.sh = nullptr;  // no smoothing lengths: the upsweep stores hmax = 0
        g.ninst = 1;     // the gravity API takes one system
        StepParams saved = h->p;
        h->p.soft2 = (float)(softening * softening);
        try {
            build_sorted(h, g, h->gpos, h->gmass, false);
            if (mode == ISPH_GRAVITY_TREE) build_tree(h, g);
            run_gravity(h, g, mode);
        } catch (...) {
            h->p = saved;
            throw;
        }
        h->p = saved;
        Dev s = g;
        s.acc = h->gpos;  // scratch for the scattered result (original order)
        launch_scatter_gravity(s, h->stream);
        std::vector<float4> out(n);
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(out.data(), h->gpos, n * sizeof(float4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            acc3[3 * i] = out[i].x;
            acc3[3 * i + 1] = out[i].y;
            acc3[3 * i + 2] = out[i].z;
        }
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_diagnostics(isph_t* h, int with_pe, double* out7) {
    return guarded(h, [&]() -> int {
        if (!out7) return fail(h, ISPH_EINVAL, "NULL output");
        if (with_pe && h->d.ninst > 1) return fail(h, ISPH_EINVAL, "the pair potential energy is per system");
        for (int k = 0; k < 7; ++k) out7[k] = 0.0;
        if (h->d.n == 0) return ISPH_OK;
        HIP_OK(hipSetDevice(h->device));
        Dev d = integ_view(h);
        launch_diag(d, h->softening, with_pe, h->stream);
        std::vector<double> red(1024 * 16);
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(red.data(), d.red, red.size() * sizeof(double), hipMemcpyDeviceToHost));
        double f[13] = {0};
        uint32_t pe_blocks = std::min<uint32_t>((d.n + 255) / 256, 1024);
        for (int b = 0; b < 1024; ++b)
            for (int k = 0; k < 12; ++k) f[k] += red[b * 16 + k];
        if (with_pe)
            for (uint32_t b = 0; b < pe_blocks; ++b) f[12] += red[b * 16 + 12];
        double M = f[2];
        out7[0] = f[0];
        out7[1] = f[1];
        out7[2] = f[12];
        // L about the COM = sum m x cross v - (sum m x / M) cross (sum m v)
        double cx = f[3] / M, cy = f[4] / M, cz = f[5] / M;
        out7[3] = f[9] - (cy * f[8] - cz * f[7]);
        out7[4] = f[10] - (cz * f[6] - cx * f[8]);
        out7[5] = f[11] - (cx * f[7] - cy * f[6]);
        out7[6] = M;
        return ISPH_OK;
    });
}

int isph_set_timing(isph_t* h, int enable) {
    return guarded(h, [&]() -> int {
        if (!enable && h->timing) flush_timing(h);
        h->timing = enable != 0;
        return ISPH_OK;
    });
}

int isph_phase_times(isph_t* h, double* ms, size_t n) {
    return guarded(h, [&]() -> int {
        if (n > 0 && !ms) return fail(h, ISPH_EINVAL, "NULL output");
        HIP_OK(hipSetDevice(h->device));
        flush_timing(h);
        for (size_t k = 0; k < n && k < (size_t)ISPH_NUM_PHASES; ++k) ms[k] = h->phase_ms[k];
        return ISPH_OK;
    });
}

// Totals of the per-wave statistics slots (sums; slots 7, 10 and 13 hold maxima).
static void read_counters(isph_t* h, unsigned long long* cnt) {
    std::vector<unsigned long long> all((size_t)kCounterSlots * kWalkWaves);
    HIP_OK(hipStreamSynchronize(h->stream));
    HIP_OK(hipMemcpy(all.data(), h->d.counters, all.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < kCounterSlots; ++k) cnt[k] = 0;
    for (uint32_t w = 0; w < kWalkWaves; ++w)
        for (uint32_t k = 0; k < kCounterSlots; ++k) {
            const unsigned long long v = all[(size_t)w * kCounterSlots + k];
            cnt[k] = (k == 7 || k == 10 || k == 13) ? std::max(cnt[k], v) : cnt[k] + v;
        }
}

int isph_kernel_stats(isph_t* h, double* out8) {
    return guarded(h, [&]() -> int {
        if (!out8) return fail(h, ISPH_EINVAL, "NULL output");
        HIP_OK(hipSetDevice(h->device));
        flush_timing(h);
        unsigned long long cnt[kCounterSlots];
        read_counters(h, cnt);
        out8[0] = h->grav_launches;
        out8[1] = h->phase_ms[ISPH_PHASE_GRAVITY];
        out8[2] = (double)cnt[0];
        out8[3] = (double)cnt[1];
        out8[4] = (double)cnt[2];
        out8[5] = (double)cnt[3];
        out8[6] = (double)cnt[4];
        out8[7] = ((double)cnt[0] + (double)cnt[1]) * kFlopsMonopole + (double)cnt[2] * kFlopsQuadrupole;
        return ISPH_OK;
    });
}

int isph_walk_stats(isph_t* h, double* out9) {
    return guarded(h, [&]() -> int {
        if (!out9) return fail(h, ISPH_EINVAL, "NULL output");
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        unsigned long long cnt[kCounterSlots];
        read_counters(h, cnt);
        for (int k = 0; k < 9; ++k) out9[k] = (double)cnt[5 + k];
        return ISPH_OK;
    });
}

int isph_clamp_energy(isph_t* h, double* joules) {
    return guarded(h, [&]() -> int {
        if (!joules) return fail(h, ISPH_EINVAL, "NULL output");
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        std::vector<double> slots(inject_slots(h->d));
        HIP_OK(hipMemcpy(slots.data(), h->d.inject, slots.size() * sizeof(double), hipMemcpyDeviceToHost));
        double sum = 0.0;
        for (double x : slots) sum += x;
        *joules = sum;
        return ISPH_OK;
    });
}

int isph_reset_stats(isph_t* h) {
    return guarded(h, [&]() -> int {
        HIP_OK(hipSetDevice(h->device));
        flush_timing(h);
        for (double& x : h->phase_ms) x = 0.0;
        h->grav_launches = 0;
        HIP_OK(hipMemsetAsync(h->d.counters, 0, (size_t)kCounterSlots * kWalkWaves * sizeof(unsigned long long),
                              h->stream));
        HIP_OK(hipMemsetAsync(h->d.inject, 0, inject_slots(h->d) * sizeof(double), h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        return ISPH_OK;
    });
}

int isph_probe_morton(isph_t* h, uint32_t* codes, float* lo3, float* hi3) {
    return guarded(h, [&]() -> int {
        if (h->d.n == 0) return ISPH_OK;
        HIP_OK(hipSetDevice(h->device));
        launch_reset_scalars(h->d, h->stream);
        launch_bbox(h->d, h->d.posh, h->stream);
        launch_morton(h->d, h->d.posh, h->stream);
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        if (codes) HIP_OK(hipMemcpy(codes, h->d.keys, (size_t)h->d.n * 4, hipMemcpyDeviceToHost));
        uint32_t k[6];
        HIP_OK(hipMemcpy(k, h->d.scal, sizeof(k), hipMemcpyDeviceToHost));
        for (int a = 0; a < 3; ++a) {
            uint32_t lo = ((k[a] >> 31) ? (k[a] ^ 0x80000000u) : ~k[a]);
            uint32_t hi = ((k[3 + a] >> 31) ? (k[3 + a] ^ 0x80000000u) : ~k[3 + a]);
            if (lo3) std::memcpy(&lo3[a], &lo, 4);
            if (hi3) std::memcpy(&hi3[a], &hi, 4);
        }
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_probe_sort(isph_t* h, uint32_t* sorted_codes, uint32_t* order) {
    return guarded(h, [&]() -> int {
        if (h->d.n == 0) return ISPH_OK;
        HIP_OK(hipSetDevice(h->device));
        build_sorted(h, h->d, h->d.posh, h->d.mass, true);
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        if (sorted_codes) HIP_OK(hipMemcpy(sorted_codes, h->d.keys, (size_t)h->d.n * 4, hipMemcpyDeviceToHost));
        if (order) HIP_OK(hipMemcpy(order, h->d.vals, (size_t)h->d.n * 4, hipMemcpyDeviceToHost));
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_probe_tree(isph_t* h, uint32_t* left, uint32_t* right, uint32_t* parent, float* com4, float* bmin4,
                    float* bmax4) {
    return guarded(h, [&]() -> int {
        if (h->d.n == 0) return 0;
        HIP_OK(hipSetDevice(h->device));
        build_sorted(h, h->d, h->d.posh, h->d.mass, true);
        build_tree(h, h->d);
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        uint32_t L = h->d.L;
        size_t nodes = 2 * (size_t)L - 1;
        if (left && L > 1) HIP_OK(hipMemcpy(left, h->d.left, (L - 1) * 4, hipMemcpyDeviceToHost));
        if (right && L > 1) HIP_OK(hipMemcpy(right, h->d.right, (L - 1) * 4, hipMemcpyDeviceToHost));
        if (parent) HIP_OK(hipMemcpy(parent, h->d.parent, nodes * 4, hipMemcpyDeviceToHost));
        if (com4) HIP_OK(hipMemcpy(com4, h->d.nA, nodes * 16, hipMemcpyDeviceToHost));
        if (bmin4)
            HIP_OK(hipMemcpy2D(bmin4, 16, &h->d.nS->lo, sizeof(SphNode), 16, nodes, hipMemcpyDeviceToHost));
        if (bmax4)
            HIP_OK(hipMemcpy2D(bmax4, 16, &h->d.nS->hi, sizeof(SphNode), 16, nodes, hipMemcpyDeviceToHost));
        h->reuse_valid = false;
        return (int)L;
    });
}

static LossParams to_loss(const isph_loss* l) {
    return LossParams{l->radius, l->r_cut, l->k_r, l->k_j, l->axis[0], l->axis[1], l->axis[2]};
}

int isph_set_state_f64(isph_t* h, const double* pos3, const double* vel3) {
    return guarded(h, [&]() -> int {
        if (h->precision == ISPH_PRECISION_FP32) return fail(h, ISPH_EINVAL, "the f64 state needs an f64 form");
        if (h->d.n == 0) return ISPH_OK;
        if (!pos3 || !vel3) return fail(h, ISPH_EINVAL, "NULL array");
        HIP_OK(hipSetDevice(h->device));
        const size_t n = h->d.n;
        StateBuf cur, b;
        try {
            cur.alloc((uint32_t)n);
            b.alloc((uint32_t)n);
            launch_save_state(h->d, cur.x, cur.v, cur.u, h->stream);  // keeps u
            HIP_OK(hipMemcpyAsync(b.x, pos3, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
            HIP_OK(hipMemcpyAsync(b.v, vel3, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
            launch_set_state(h->d, b.x, b.v, cur.u, h->stream);
            HIP_OK(hipStreamSynchronize(h->stream));
            HIP_OK(hipGetLastError());
        } catch (...) {
            cur.release();
            b.release();
            throw;
        }
        cur.release();
        b.release();
        h->reuse_valid = false;
        return ISPH_OK;
    });
}

int isph_state_f64(isph_t* h, double* pos3, double* vel3) {
    return guarded(h, [&]() -> int {
        if (h->precision == ISPH_PRECISION_FP32) return fail(h, ISPH_EINVAL, "the f64 state needs an f64 form");
        if (h->d.n == 0) return ISPH_OK;
        HIP_OK(hipSetDevice(h->device));
        HIP_OK(hipStreamSynchronize(h->stream));
        const size_t n = h->d.n;
        if (pos3) HIP_OK(hipMemcpy(pos3, h->d.pos64, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
        if (vel3) HIP_OK(hipMemcpy(vel3, h->d.vel64, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
        return ISPH_OK;
    });
}

int isph_loss_value(isph_t* h, const isph_loss* loss, double* value) {
    return guarded(h, [&]() -> int {
        if (!loss || !value) return fail(h, ISPH_EINVAL, "NULL argument");
        if (h->precision == ISPH_PRECISION_FP32) return fail(h, ISPH_EINVAL, "the loss needs an f64 form");
        if (h->d.n == 0) {
            *value = 0.0;
            return ISPH_OK;
        }
        HIP_OK(hipSetDevice(h->device));
        ensure_adj(h);
        launch_adj_loss(h->d, h->adj, to_loss(loss), 1.0, h->stream);
        HIP_OK(hipMemcpyAsync(value, h->adj.red + 7, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipGetLastError());
        return ISPH_OK;
    });
}

int isph_vjp_run(isph_t* h, float cfl, const float* dts_in, uint32_t steps, uint32_t checkpoint_every,
                 const isph_loss* loss, double* loss_out, double* dpos3, double* dvel3, double* du, double* dmass,
                 double* deos4, float* dts_out, double* ms2) {
    return guarded(h, [&]() -> int {
        if (!loss) return fail(h, ISPH_EINVAL, "NULL loss");
        if (h->precision != ISPH_PRECISION_MIXED) return fail(h, ISPH_EINVAL, "the adjoint needs the mixed form");
        if (h->d.ninst > 1) return fail(h, ISPH_EINVAL, "the adjoint takes one system, not an ensemble");
        if (h->p.bulk_cr.w != 0.0f || h->p.n_ext != 0)
            return fail(h, ISPH_EINVAL, "the adjoint does not differentiate the bulk or external masses");
        if (h->gravity_mode == ISPH_GRAVITY_TREE && (h->p.quad || h->p.grav_ppl != 1))
            return fail(h, ISPH_EINVAL, "the adjoint's tree gravity is the monopole walk with one particle per lane");
        if (steps == 0 || checkpoint_every == 0) return fail(h, ISPH_EINVAL, "steps and checkpoint_every must be > 0");
        if (!dts_in && !(cfl > 0.0f)) return fail(h, ISPH_EINVAL, "cfl must be > 0 without a dt sequence");
        if (h->d.n == 0) return fail(h, ISPH_EINVAL, "no particles");
        HIP_OK(hipSetDevice(h->device));
        ensure_adj(h);
        const uint32_t n = h->d.n, T = steps, K = std::min(checkpoint_every, steps);
        const uint32_t nseg = (T + K - 1) / K;
        const LossParams lp = to_loss(loss);
        VjpScratch w;
        w.ck.resize(nseg);
        for (auto& b : w.ck) b.alloc(n);
        w.start.resize(K);
        w.mid.resize(K);
        for (auto& b : w.start) b.alloc(n);
        for (auto& b : w.mid) b.alloc(n);
        w.ev.resize(2 * K);
        for (auto& e : w.ev) e.alloc(h->d.cap, h->d.nsort_blocks);
        w.final_state.alloc(n);
        dalloc(&w.dts, T);
        HIP_OK(hipStreamSynchronize(h->stream));
        const bool saved_reuse = h->reuse_enabled;
        h->reuse_enabled = false;
        h->reuse_valid = false;
        hipEvent_t e0, e1, e2;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        HIP_OK(hipEventCreate(&e2));
        struct EvGuard {
            hipEvent_t* e;
            ~EvGuard() {
                for (int k = 0; k < 3; ++k) (void)hipEventDestroy(e[k]);
            }
        };
        hipEvent_t evs[3] = {e0, e1, e2};
        EvGuard eg{evs};
        struct ReuseGuard {
            isph_ctx* c;
            bool v;
            ~ReuseGuard() { c->reuse_enabled = v; c->reuse_valid = false; }
        } rg{h, saved_reuse};
        const Dev& d = h->d;
        const hipStream_t st = h->stream;

        // forward, checkpointing every K steps; the device records each step's dt
        HIP_OK(hipEventRecord(e0, st));
        for (uint32_t s = 0; s < T; ++s) {
            if (s % K == 0) launch_save_state(d, w.ck[s / K].x, w.ck[s / K].v, w.ck[s / K].u, st);
            if (dts_in) {
                launch_set_dt_value(d, dts_in[s], st);
                kdk(h, false, 0.0f);
            } else {
                kdk(h, true, cfl);
            }
            HIP_OK(hipMemcpyAsync(w.dts + s, d.dt, sizeof(float), hipMemcpyDeviceToDevice, st));
        }
        HIP_OK(hipEventRecord(e1, st));
        launch_save_state(d, w.final_state.x, w.final_state.v, w.final_state.u, st);
        const double scale = 1.0;
        launch_adj_loss(d, h->adj, lp, scale, st);
        HIP_OK(hipMemsetAsync(h->adj.pbar, 0, kMaxMaterials * 4 * sizeof(double), st));
        std::vector<float> dts(T);
        HIP_OK(hipMemcpyAsync(dts.data(), w.dts, T * sizeof(float), hipMemcpyDeviceToHost, st));
        double lval = 0.0;
        HIP_OK(hipMemcpyAsync(&lval, h->adj.red + 7, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));

        // reverse: recompute each segment from its checkpoint (storing every evaluation), then walk back
        for (int seg = (int)nseg - 1; seg >= 0; --seg) {
            const uint32_t s0 = (uint32_t)seg * K, s1 = std::min(T, s0 + K);
            launch_set_state(d, w.ck[seg].x, w.ck[seg].v, w.ck[seg].u, st);
            for (uint32_t s = s0; s < s1; ++s) {
                const uint32_t i = s - s0;
                launch_save_state(d, w.start[i].x, w.start[i].v, w.start[i].u, st);
                launch_set_dt_value(d, dts[s], st);
                eval_into(h, w.ev[2 * i]);
                launch_kick_drift(d, h->p, st);
                launch_save_state(d, w.mid[i].x, w.mid[i].v, w.mid[i].u, st);
                eval_into(h, w.ev[2 * i + 1]);
                launch_kick(d, st);
            }
            for (int s = (int)s1 - 1; s >= (int)s0; --s) {
                const uint32_t i = (uint32_t)s - s0;
                launch_vjp_kick(d, h->adj, w.mid[i].u, w.ev[2 * i + 1].acc, dts[s], st);
                vjp_of(h, w.ev[2 * i + 1]);
                launch_vjp_kick_drift(d, h->adj, w.start[i].u, w.ev[2 * i].acc, dts[s], st);
                vjp_of(h, w.ev[2 * i]);
            }
        }
        HIP_OK(hipEventRecord(e2, st));
        // leave the handle at the forward run's final state
        launch_set_state(d, w.final_state.x, w.final_state.v, w.final_state.u, st);
        HIP_OK(hipStreamSynchronize(st));
        HIP_OK(hipGetLastError());
        if (ms2) {
            float a = 0.f, b = 0.f;
            HIP_OK(hipEventElapsedTime(&a, e0, e1));
            HIP_OK(hipEventElapsedTime(&b, e1, e2));
            ms2[0] = a;
            ms2[1] = b;
        }
        if (loss_out) *loss_out = lval;
        auto fetch = [&](double* dst, const double* src, size_t count) {
            if (!dst) return;
            HIP_OK(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < count; ++k) dst[k] /= scale;
        };
        fetch(dpos3, h->adj.xb, 3 * (size_t)n);
        fetch(dvel3, h->adj.vb, 3 * (size_t)n);
        fetch(du, h->adj.ub, n);
        fetch(dmass, h->adj.mb, n);
        fetch(deos4, h->adj.pbar, 4 * (size_t)kMaxMaterials);
        if (dts_out) std::copy(dts.begin(), dts.end(), dts_out);
        return ISPH_OK;
    });
}

}  // extern "C"

// ---- autotuner --------------------------------------------------------------------------------------------
namespace {

struct Tuning {
    uint32_t K;
    float extent;
    int ppl, expand;
};

// Snapshot of the integrated state, restored after each trial.
struct Snapshot {
    float4 *posh = nullptr, *velu = nullptr;
    double *pos64 = nullptr, *vel64 = nullptr, *itime = nullptr, time = 0.0;
    void take(isph_ctx* c) {
        const Dev& d = c->d;
        const size_t n = d.n;
        dalloc(&posh, n);
        dalloc(&velu, n);
        HIP_OK(hipMemcpyAsync(posh, d.posh, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(velu, d.velu, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        if (d.pos64 && c->precision != ISPH_PRECISION_FP32) {
            dalloc(&pos64, 3 * n);
            dalloc(&vel64, 3 * n);
            HIP_OK(hipMemcpyAsync(pos64, d.pos64, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            HIP_OK(hipMemcpyAsync(vel64, d.vel64, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        }
        if (d.ninst > 1) {
            dalloc(&itime, d.ninst);
            HIP_OK(hipMemcpyAsync(itime, d.itime, d.ninst * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        }
        HIP_OK(hipMemcpyAsync(&time, d.time, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    void restore(isph_ctx* c) {
        const Dev& d = c->d;
        const size_t n = d.n;
        HIP_OK(hipMemcpyAsync(d.posh, posh, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(d.velu, velu, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        if (pos64) {
            HIP_OK(hipMemcpyAsync(d.pos64, pos64, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            HIP_OK(hipMemcpyAsync(d.vel64, vel64, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        }
        if (itime)
            HIP_OK(hipMemcpyAsync(d.itime, itime, d.ninst * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(d.time, &time, sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->reuse_valid = false;
    }
    ~Snapshot() {
        void* ptrs[] = {posh, velu, pos64, vel64, itime};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
    }
};

// A shape read from the cache file is taken only if isph_set_gravity and isph_set_walk would accept it for this
// handle: the file is outside the library's control, and a leaf size of 0 or above 64 would break the tree.
bool valid_tuning(const isph_ctx* c, const Tuning& t) {
    return t.K >= 1 && t.K <= 64 && (c->d.ninst <= 1 || t.K == 1) && std::isfinite(t.extent) && t.extent > 0.0f &&
           (t.ppl == 1 || t.ppl == 2) && t.expand >= 1 && t.expand <= 64;
}

void apply(isph_ctx* c, const Tuning& t) {
    drop_graph(c);
    c->d.K = t.K;
    set_leaves(c, c->d.n);
    c->p.group_extent = t.extent;
    c->p.grav_ppl = t.ppl;
    c->p.sph_expand = t.expand;
    c->reuse_valid = false;
}

std::string tune_key(isph_ctx* c) {
    uint32_t bucket = 0;
    while ((1u << (bucket + 1)) <= c->d.n) ++bucket;
    std::ostringstream k;
    k << "n2^" << bucket << " inst" << (c->d.ninst > 1 ? 1 : 0) << " prec" << c->precision << " grav"
      << c->gravity_mode << " quad" << c->p.quad;
    return k.str();
}

}  // namespace

extern "C" {

int isph_autotune(isph_t* h, float cfl, uint32_t trial_steps, const char* cache_path, isph_tuning* chosen) {
    return guarded(h, [&]() -> int {
        if (!(cfl > 0.0f) || trial_steps == 0) return fail(h, ISPH_EINVAL, "cfl and trial_steps must be > 0");
        if (h->d.n == 0) return fail(h, ISPH_EINVAL, "no particles");
        HIP_OK(hipSetDevice(h->device));
        const std::string key = tune_key(h);
        if (cache_path) {  // "key | K extent ppl expand ms" lines
            std::ifstream in(cache_path);
            std::string line;
            while (std::getline(in, line)) {
                const size_t bar = line.find(" | ");
                if (bar == std::string::npos || line.substr(0, bar) != key) continue;
                std::istringstream v(line.substr(bar + 3));
                Tuning t{};
                double ms = 0.0;
                if (!(v >> t.K >> t.extent >> t.ppl >> t.expand >> ms) || !valid_tuning(h, t)) continue;
                apply(h, t);
                if (chosen) *chosen = isph_tuning{t.K, t.extent, t.ppl, t.expand, ms};
                return 1;
            }
        }
        Snapshot snap;
        snap.take(h);
        const bool graph = h->graph_on;
        h->graph_on = false;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        struct EvGuard {
            hipEvent_t a, b;
            ~EvGuard() {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        } eg{e0, e1};
        auto trial = [&](const Tuning& t) -> double {
            apply(h, t);
            step_once(h, true, cfl);  // warm-up step (group lists, caches)
            HIP_OK(hipEventRecord(e0, h->stream));
            for (uint32_t s = 0; s < trial_steps; ++s) step_once(h, true, cfl);
            HIP_OK(hipEventRecord(e1, h->stream));
            HIP_OK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            snap.restore(h);
            return ms / trial_steps;
        };
        // Coordinate descent from the current shape: leaf size (one system only), group extent, gravity
        // particles per lane, SPH whole-subtree size.
        Tuning best{h->d.K, h->p.group_extent, h->p.grav_ppl, h->p.sph_expand};
        double best_ms = trial(best);
        auto try_axis = [&](auto setter, std::initializer_list<double> values) {
            Tuning base = best;
            for (double v : values) {
                Tuning t = base;
                setter(t, v);
                if (t.K == best.K && t.extent == best.extent && t.ppl == best.ppl && t.expand == best.expand) continue;
                double ms = trial(t);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = t;
                }
            }
        };
        if (h->d.ninst == 1 && h->gravity_mode == ISPH_GRAVITY_TREE)
            try_axis([](Tuning& t, double v) { t.K = (uint32_t)v; }, {1, 2, 4});
        try_axis([](Tuning& t, double v) { t.extent = (float)v; }, {3.0, 4.0, 6.0});
        if (h->gravity_mode == ISPH_GRAVITY_TREE) try_axis([](Tuning& t, double v) { t.ppl = (int)v; }, {1, 2});
        try_axis([](Tuning& t, double v) { t.expand = (int)v; }, {1, 4, 8});
        apply(h, best);
        h->graph_on = graph;
        if (cache_path) {
            std::ofstream out(cache_path, std::ios::app);
            out << key << " | " << best.K << " " << best.extent << " " << best.ppl << " " << best.expand << " "
                << best_ms << "\n";
        }
        if (chosen) *chosen = isph_tuning{best.K, best.extent, best.ppl, best.expand, best_ms};
        return 0;
    });
}

}  // extern "C"
