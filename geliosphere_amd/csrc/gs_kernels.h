// gs_kernels.h - host-visible interface of the trajectory kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/gs_api.h"

namespace gs {
// Proton mass, charge and the speed of light as the reference writes them
// (CpuImplementations/include/Constants.hpp:63-87), and the rest energy T0.
constexpr double kM0 = 1.67261e-27;
constexpr double kQ = 1.60219e-19;
constexpr double kC = 2.99793e8;
constexpr double kT0 = kM0 * kC * kC / (kQ * 1e9);  // [GeV]
constexpr double kRh = 695510.0 / 150000000.0;      // solar radius [AU] (Constants.hpp:127)
constexpr double kGeC = (5.0 / 3.4) / 3.0;           // K_par factor (GeliosphereCpuModel.cpp:167)
constexpr double kGeKf = 1.0 / (3.0 * (5.0 / 3.4));  // derivative factor (GeliosphereCpuModel.cpp:210-299)

constexpr int kWave = 64;
constexpr int kBlockThreads = 256;  // default workgroup size (trajectory kernels: 128, 256 or 512)
constexpr int kCycleSteps = 4;     // steps between refills (whole Philox blocks for every model)
constexpr uint64_t kChunk = 64;    // default particle ids a wave takes from the global queue per grab
constexpr int kBatchMaxSims = 1024; // simulations per batch trajectory launch (offsets in LDS)

// Indices of the statistics words (LDS and global).
enum Stat {
    ST_PARTICLES = 0,
    ST_EVENTS,
    ST_OUTSIDE,
    ST_STEPS,
    ST_NAN,
    ST_CAP,
    ST_INEXACT,
    ST_OVERFLOW,
    ST_CYCLES,    // wave cycles issued (x 64 x kCycleSteps = lane-step slots)
    ST_NREC,      // events offered to the record buffer
    ST_MAX_STEPS, // plain runs: most steps of one particle (a maximum, not a sum)
    ST_COUNT_
};

// Parameters a model's init/step read (host-derived constants included).
// Batched runs copy them per lane into registers; plain runs read them
// wave-uniformly.
struct StepParams {
    double K0, V, dt, r_inj, theta_inj, tilt, polarity, ratio, delta0;
    double sqdt;          // sqrt(dt)
    double two_dt;        // 2 dt
    double sk2dt;         // sqrt(2 K0 dt T0) (the 1D models carry momenta in units of T0)
    double vdt;           // signed V dt as the model uses it (1D B-p: -V dt)
    double cp;            // (2/3) V dt with the same sign (momentum / energy loss)
    double csum;          // (4/3) V dt (1D F-p loss integral)
    double omega_v;       // omega / V
    double omega2_v2;     // omega^2 / V^2
    double th_lo, th_hi;  // 1.7 deg and 178.3 deg in rad, as the reference writes them
    double pol_theta;     // polarity used in the 2D theta drift
    // 2D products of per-simulation and fixed factors (the step multiplies
    // each once; with C = (5/3.4)/3 and kf = 1/(3 (5/3.4)), the factors
    // GeliosphereCpuModel.cpp:167-299 and SolarPropLikeCpuModel.cpp:131-163
    // write around K0 and K0 beta P):
    double ge_k0c;        // K0 C (both 2D models)
    double ge_a1;         // ratio kf / C
    double ge_a2;         // (1 - ratio) kf / C
    double ge_a3;         // 3 (1 - ratio) / C
    double poldt;         // polarity dt (2D drifts)
    double polthdt;       // pol_theta dt (2D SolarProp theta drift)
    double sq2dt;         // sqrt(2 dt)
    double delta0_rh;     // delta0 / rh
};

// Scheduling order of a run's particles (gs_launch_order, batches): offsets
// sorted by a bucket of their injection energy, bucket = (bits(T) >> shift)
// - kmin clamped to [0, nb); the bits of a positive double grow with its
// value.  nb = 0: id order.  nb <= kOrderSimBuckets.
constexpr uint32_t kOrderMaxBuckets = 1024;
constexpr uint32_t kOrderSimBuckets = 512;
struct OrderKey {
    uint32_t shift, nb;
    uint64_t kmin;
};

// Parameters of one simulation.
struct DevParams : StepParams {
    OrderKey order;       // batches: this simulation's energy buckets
    int32_t model, injection;
    uint32_t flags, n_bins, n_weights, grid_n;
    int32_t weights[GS_MAX_WEIGHTS];
    double T_min, T_max;
    double log_ratio;     // ln(T_max / T_min)
    double grid_step;     // (T_max - T_min) / grid_n
    uint64_t n_particles, first_particle, seed, max_steps;
    double edges[GS_MAX_BINS + 1];
};

// One event as the trajectory kernels store it (exit or 1 AU crossing).
// One whole 64-byte line per record: aligned full-line stores need no
// read-modify-write in L2.
struct alignas(64) DevEvent {
    double t_inj, t_exit, r, theta, sumac;
    uint64_t id;
    uint32_t sim;   // simulation index within its launch (0 for plain runs)
    uint32_t pad;
    uint64_t pad2;
};
static_assert(sizeof(DevEvent) == 64, "DevEvent is one 64-byte line");

// A live particle handed from one plain launch to the next (see DevOut):
// the model's state, its id, steps taken and next Philox block.  The
// counter-based normal stream makes the continuation exact.
struct alignas(64) SpillRec {
    double s[6];
    uint64_t pid;
    uint32_t steps, blk;
};
static_assert(sizeof(SpillRec) == 64, "SpillRec is one 64-byte line");
constexpr int kSpillLanes = 16;      // a wave at or below this many live lanes hands them on
constexpr uint32_t kSpillMinCycles = 4;   // after at least this many cycles in its launch

// Device buffers of a launch over n_sims simulations (plain runs: 1).
// Per simulation s: histogram count[GS_MAX_BINS], digits[GS_MAX_WEIGHTS][2]
// [GS_MAX_BINS][GS_ACC_LIMBS] (32-bit values in 64-bit words), stats
// [ST_COUNT_], and an event region events[ev_off[s] .. ev_off[s]+ev_cap[s]).
struct DevOut {
    uint64_t *queue;          // virtual particle ids handed out
    uint64_t *g_count;
    uint64_t *g_digits;
    uint64_t *g_stats;
    DevEvent *events;
    const uint64_t *ev_off;
    const uint64_t *ev_cap;
    uint64_t *n_events;       // per simulation; may exceed ev_cap (then re-run with more room)
    uint64_t *cycles;         // wave cycles issued by the launch
    const uint32_t *order;    // queue position (plain) or virtual id (batch) -> particle offset
                              // within its simulation (nullptr: identity)
    // Plain runs only.  Tail compaction: once the queue is exhausted, a wave
    // with at most kSpillLanes live lanes writes them to spill[*spill_count++]
    // and exits (spill == nullptr: never).  Slots at or beyond spill_cap are
    // not written: such particles go on in place.  A continuation launch
    // takes its particles from resume[0 .. resume_n) instead of the queue of
    // new ids.
    SpillRec *spill;
    uint64_t *spill_count;
    uint64_t spill_cap;
    const SpillRec *resume;
    uint64_t resume_n;
    // Particle ids a wave takes from the queue per grab (0: kChunk): larger
    // grabs mean fewer queue atomics, smaller ones a finer share-out of the
    // last particles among the waves.
    uint32_t chunk;
};

// LDS layout per workgroup: count[n_bins] | digits[n_weights][2][n_bins][16] | stat[ST_COUNT_]
__host__ __device__ inline int lds_words(int nb, int nw) {
    return nb + nw * 2 * nb * GS_ACC_LIMBS + ST_COUNT_;
}

constexpr size_t kDigitsPerSim = (size_t)GS_MAX_WEIGHTS * 2 * GS_MAX_BINS * GS_ACC_LIMBS;

// Fixed device scratch of the no-allocation plain run (gs_run_plain_noalloc),
// in 64-bit words: one simulation's histogram words, its parameters and run
// counters (kScratchHeadWords), the energy-order counters and permutation of
// up to kScratchEvents particles, and an event region of kScratchEvents
// 64-byte records.  34.2 MiB per device.
constexpr uint64_t kScratchEvents = 1ull << 19;
constexpr size_t kScratchHeadWords = 8704;
constexpr size_t kScratchCountWords = kOrderMaxBuckets;   // 2 x kOrderMaxBuckets uint32
constexpr size_t kScratchPermWords = kScratchEvents / 2;  // kScratchEvents uint32
constexpr size_t kScratchWords =
    kScratchHeadWords + kScratchCountWords + kScratchPermWords + kScratchEvents * (sizeof(DevEvent) / 8);
}  // namespace gs

// Device address and size of `device`'s scratch array.  The first call for a
// device must be made with that device current; the address is then kept
// per device.
hipError_t gs_scratch_address(int device, void **ptr, size_t *bytes);

hipError_t gs_launch_zero(uint64_t *p, size_t n, hipStream_t s);
hipError_t gs_launch_debug(int fn, const double *x, double *y, uint64_t n, uint64_t seed,
                           uint64_t pid, hipStream_t s);
// Workgroup sizes the trajectory kernels are built for.
bool gs_traj_block_supported(unsigned block_threads);
// Blocks of `block_threads` a CU holds for the trajectory kernel of `model`.
hipError_t gs_traj_occupancy(int model, bool batch, unsigned block_threads, int *blocks_per_cu);
// Trajectories of n_sims simulations of one model.  sims: device array;
// offsets: device array of n_sims + 1 prefix sums of n_particles (the
// virtual id space).  batch = false requires n_sims == 1 (uniform parameters).
hipError_t gs_launch_traj(int model, bool batch, const gs::DevParams *sims, uint32_t n_sims,
                          const uint64_t *offsets, const gs::DevOut &O, unsigned grid,
                          unsigned block_threads, hipStream_t s);
// Scheduling order of the n = sims->n_particles particles of one plain
// simulation: perm[0 .. n) receives the offsets 0 .. n-1 in ascending bucket
// order (within a bucket in no fixed order).  counts: 2 * key.nb words of
// scratch.
hipError_t gs_launch_order(const gs::DevParams *sims, uint64_t n, const gs::OrderKey &key, uint32_t *counts,
                           uint32_t *perm, hipStream_t s);
// Scheduling order of a batch launch: for the n_sims simulations of the
// virtual id space `offsets` (device, n_sims + 1 prefix sums), perm[v]
// receives a particle offset within v's simulation such that each
// simulation's range is in ascending energy-bucket order (sims[s].order; id
// order where nb = 0).  counts: 2 * n_sims * kOrderSimBuckets words.
hipError_t gs_launch_batch_order(const gs::DevParams *sims, uint32_t n_sims, const uint64_t *offsets, uint64_t n,
                                 uint32_t *counts, uint32_t *perm, hipStream_t s);
// Histograms of n_sims simulations from their event regions; lds_bytes fits
// the largest (n_bins, n_weights) among them.
hipError_t gs_launch_fold(const gs::DevParams *sims, uint32_t n_sims, const gs::DevOut &O,
                          size_t lds_bytes, unsigned blocks_per_sim, hipStream_t s);
