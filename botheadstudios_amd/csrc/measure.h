// On-device measurement of an impact: the disk view (gpu_sph.rs:689-785 with orbit.rs:91-117) and the
// friends-of-friends clump finder with per-clump classification (accretion.rs:158-301).
//
// Every reduction runs over fixed chunks of at most kMeasChunk consecutive particles that never span two
// instances, summed in a fixed order, so each result is reproducible bit for bit from run to run.
#pragma once

#include "state.h"

namespace isph {

constexpr uint32_t kMeasChunk = 1024;  // particles per reduction chunk (256 threads x 4)
constexpr int kPartFields = 16;        // per-chunk partial sums
constexpr int kViewFields = 24;        // per-instance view record (doubles)
constexpr int kClumpFields = 24;       // per-clump record (doubles)
constexpr int kParamFields = 12;       // per-instance clump-finder parameters (doubles)

// view record fields
enum : int {
    kVValid = 0,
    kVNLive = 1,
    kVCom = 2,   // 3
    kVVCom = 5,  // 3
    kVRRem = 8,
    kVMRem = 9,
    kVMu = 10,
    kVDisk0 = 11,
    kVDiskO = 12,
    kVEsc = 13,
    kVMeanH = 14,
    kVRem0 = 15,
    kVRemO = 16,
    kVEsc0 = 17,
    kVNDisk = 18,
    kVMTot = 19,
};
// clump record fields
enum : int {
    kCId = 0,  // first member slot (global); made instance-local on the host
    kCN = 1,
    kCMass = 2,
    kCCom = 3,   // 3
    kCVel = 6,   // 3
    kCRho = 9,
    kCRadius = 10,
    kCKe = 11,
    kCPe = 12,
    kCL = 13,    // 3
    kCThermal = 16,
    kCBound = 17,
    kCOutside = 18,
    kCInst = 19,
};
// clump-finder parameters
enum : int { kPLink = 0, kPSoft2 = 1, kPG = 2, kPCPos = 3, kPCMass = 6, kPCRad = 7 };

struct Meas {
    uint32_t cap = 0, cap_chunks = 0, cap_inst = 0;
    // chunking of [0, n): (start, end), the instance of each chunk, the first chunk of each instance
    uint2* chunks = nullptr;
    uint32_t* cinst = nullptr;
    uint32_t* ichunk = nullptr;
    uint32_t nchunks = 0;
    double* part = nullptr;   // [cap_chunks * kPartFields]
    double* coff = nullptr;   // [cap_chunks] cumulative mass before each chunk (radius order)
    double* view = nullptr;   // [cap_inst * kViewFields]
    uint32_t* cross = nullptr;  // [cap_inst] radius-sorted position where the enclosed mass reaches 85 %
    // sort buffers (radius sort, cell-hash sort, label sort)
    uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *khi = nullptr;
    double* scum = nullptr;   // [cap] cumulative mass within the chunk, radius order
    uint32_t* slot = nullptr; // [cap] a disk particle's slot within its chunk, or ~0
    uint32_t* mbase = nullptr;  // [cap_chunks] member base of each chunk
    uint32_t* moff = nullptr;   // [cap_inst + 1] first member of each instance
    // members (f64): (x, y, z, m), (vx, vy, vz, rho); instance; index within the instance's particles
    double4* mp = nullptr;
    double4* mv = nullptr;
    uint32_t* minst = nullptr;
    uint32_t* midx = nullptr;
    double* iparam = nullptr;  // [cap_inst * kParamFields]
    // friends-of-friends
    uint32_t* parent = nullptr;
    uint32_t* label = nullptr;
    // clumps
    uint32_t* cflag = nullptr;   // [cap + 1] scan buffer
    uint32_t* cstart = nullptr;  // [cap + 1]
    uint32_t* tcnt = nullptr;    // [cap + 1] PE tiles per clump, scanned in place into offsets
    double4* cp = nullptr;       // members in clump order
    double4* cv = nullptr;
    double* crec = nullptr;      // [cap * kClumpFields]
    double* tpe = nullptr;       // per-tile partial pair sums
    size_t tpe_cap = 0;
    uint32_t* word = nullptr;    // [24] device words: [0] members; [8] clumps, [9] union overflow;
                                 //      [16] selected clumps
    uint32_t *ks = nullptr, *vs = nullptr;  // members sorted by clump label: labels, member slots
    unsigned long long* csum = nullptr;  // [cap_inst * 3] clumps >= 2 members, self-bound, moon mass bits
    uint32_t* sel = nullptr;     // [cap] selected clump indices
    double* srec = nullptr;      // [cap * kClumpFields] selected records
    // state of the last measurement
    uint32_t n_members = 0, n_clumps = 0, ninst = 0;
    bool valid = false;
};

// Launchers (all on stream s). The host reads m.word between the stages to size the next launches.
// Live sums per chunk and per instance (COM, bulk velocity, mean h, validity).
void launch_meas_live(const Dev& d, Meas& m, uint32_t ninst, hipStream_t s);
// Radius sort per instance and the 85 %-mass remnant (radius, enclosed mass, mu).
void launch_meas_remnant(const Dev& d, Meas& m, uint32_t ninst, hipStream_t s);
// Bound / perigee classification, per-instance sums, disk members gathered in upload order
// (word[0] = member count) and the clump-finder parameters of each instance.
void launch_meas_classify(const Dev& d, Meas& m, uint32_t ninst, double clump_soft, double link_factor,
                          hipStream_t s);
// Friends-of-friends over members [0, n) and the clump segments (word[8] = clump count, word[9] != 0 if a
// union did not settle).
void launch_fof(Meas& m, uint32_t n, uint32_t* hist, hipStream_t s);
// Per-clump statistics and the self-energy tile offsets (tcnt[n_clumps] = tile count).
void launch_clump_stats(Meas& m, uint32_t n_clumps, hipStream_t s);
// Exact softened self potential energy over the tiles, then the bound / Roche gates.
void launch_clump_pe(Meas& m, uint32_t n_clumps, uint32_t n_tiles, hipStream_t s);
// Indices of instance `inst`'s clumps with >= min_members members into m.sel (word[16] = count), then
// their records into m.srec.
void launch_clump_select(Meas& m, uint32_t n_clumps, uint32_t inst, uint32_t min_members, hipStream_t s);
void launch_clump_gather(Meas& m, uint32_t count, hipStream_t s);
// Per-instance clump counts (>= 2 members, self-bound) and the largest accreting clump's mass into m.csum.
void launch_clump_summary(Meas& m, uint32_t n_clumps, uint32_t ninst, hipStream_t s);

}  // namespace isph
