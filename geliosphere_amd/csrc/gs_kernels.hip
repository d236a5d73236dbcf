// gs_kernels.hip - trajectory kernels for gfx950.
//
// Work distribution: a persistent grid of 256-thread workgroups (4 waves).
// Each wave pulls particle ids from a global queue in chunks of 64, keeps a
// private window of ids and, at every cycle boundary (kCycleSteps steps),
// hands new particles to the lanes whose particle has ended.  A cycle always
// consumes whole Philox blocks, so at a cycle boundary every lane's normal
// stream is block-aligned and the RNG branches stay wave-uniform.
//
// The reference launches 4096x256 threads per iteration with one particle
// per thread and waits for the slowest particle of every launch
// (CUDAKernel/src/OneDimensionBpModel.cu:171-184); here no lane waits for
// another lane's particle except at the very end of the run.
#include <algorithm>
#include <mutex>

#include "gs_device.hpp"
#include "gs_kernels.h"

namespace gs {

// Append one event to its simulation's event region (counted even when the
// region is full; the host then runs the simulation again with more room).
__device__ __forceinline__ void emit_event(const DevOut &O, uint64_t pid, uint32_t sim, double t_inj, double t_exit,
                                        double r, double theta, double sumac) {
    // The lanes emitting together nearly always share the simulation: then
    // the first of them reserves the slots of all with one atomic (one
    // counter per simulation takes every exit of a 1D batch otherwise).
    const uint64_t lanes = __ballot(1);
    const uint32_t s0 = __builtin_amdgcn_readfirstlane(sim);
    unsigned long long i;
    if (__ballot(sim != s0) == 0) {
        const int lane = (int)(threadIdx.x & (kWave - 1));
        const int leader = __ffsll((unsigned long long)lanes) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd((unsigned long long *)&O.n_events[s0], (unsigned long long)__popcll(lanes));
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)base, leader);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(base >> 32), leader);
        i = (((unsigned long long)hi << 32) | lo) + (unsigned long long)__popcll(lanes & ((1ull << lane) - 1));
    } else {
        i = atomicAdd((unsigned long long *)&O.n_events[sim], 1ull);
    }
    if (i < O.ev_cap[sim]) {
        // Streaming stores: a record is written once and read by the fold
        // pass, so it need not occupy L2 (and a lone record does not make
        // the L2 fetch the rest of its line).
        DevEvent *d = &O.events[O.ev_off[sim] + i];
        __builtin_nontemporal_store(t_inj, &d->t_inj);
        __builtin_nontemporal_store(t_exit, &d->t_exit);
        __builtin_nontemporal_store(r, &d->r);
        __builtin_nontemporal_store(theta, &d->theta);
        __builtin_nontemporal_store(sumac, &d->sumac);
        __builtin_nontemporal_store(pid, &d->id);
        __builtin_nontemporal_store((uint64_t)sim, (uint64_t *)&d->sim);
        __builtin_nontemporal_store((uint64_t)0, &d->pad2);
    }
}

// Event staging: each wave appends its records to its own
// LDS area (slot from an LDS counter) and, at a cycle boundary once at least
// kStageFlush are staged, writes them out together: the wave reserves a
// contiguous range of the event region and its 64 lanes store consecutive
// 8-byte words, so records reach HBM as full lines instead of four partial
// 16-byte writes from one lane.  Records that find the area full are written
// directly.
constexpr int kStage = 32;         // records per wave
constexpr int kStageFlush = 16;
constexpr int kRecWords = (int)(sizeof(DevEvent) / sizeof(uint64_t));

template <int kWaves>
struct Stage {
    uint64_t words[kWaves][kStage * kRecWords];
    uint32_t count[kWaves];
};

// `sim` is the record's simulation within the launch (0 for plain runs).
template <class St>
__device__ __forceinline__ void stage_event(St *st, const DevOut &O, uint64_t pid, uint32_t sim, double t_inj,
                                            double t_exit, double r, double theta, double sumac) {
    const int w = (int)(threadIdx.x / kWave);
    const uint32_t slot = atomicAdd(&st->count[w], 1u);
    if (slot < (uint32_t)kStage) {
        uint64_t *d = &st->words[w][slot * kRecWords];
        d[0] = fm::double_to_bits(t_inj);
        d[1] = fm::double_to_bits(t_exit);
        d[2] = fm::double_to_bits(r);
        d[3] = fm::double_to_bits(theta);
        d[4] = fm::double_to_bits(sumac);
        d[5] = pid;
        d[6] = sim;
        d[7] = 0;
    } else {
        emit_event(O, pid, sim, t_inj, t_exit, r, theta, sumac);
    }
}

// Called by the whole wave (uniform control flow).  In a batch the staged
// records normally share one simulation: they then go out as one range of
// its region; otherwise each goes out on its own.
template <bool kBatch, class St>
__device__ __forceinline__ void stage_flush(St *st, const DevOut &O, uint32_t min_count) {
    const int w = (int)(threadIdx.x / kWave), lane = (int)(threadIdx.x & (kWave - 1));
    __builtin_amdgcn_wave_barrier();
    uint32_t c = st->count[w];
    if (c > (uint32_t)kStage) c = kStage;
    if (c < min_count || c == 0) return;
    uint32_t sim = 0;
    if constexpr (kBatch) {
        sim = (uint32_t)st->words[w][6];   // record 0's simulation
        const uint32_t mine = (uint32_t)lane < c ? (uint32_t)st->words[w][lane * kRecWords + 6] : sim;
        if (__ballot(mine != sim) != 0) {
            if ((uint32_t)lane < c) {
                const uint64_t *d = &st->words[w][lane * kRecWords];
                emit_event(O, d[5], mine, fm::bits_to_double(d[0]), fm::bits_to_double(d[1]),
                           fm::bits_to_double(d[2]), fm::bits_to_double(d[3]), fm::bits_to_double(d[4]));
            }
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) st->count[w] = 0;
            __builtin_amdgcn_wave_barrier();
            return;
        }
    }
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd((unsigned long long *)&O.n_events[sim], (unsigned long long)c);
    base = __shfl(base, 0);
    const uint64_t cap = O.ev_cap[sim];
    uint64_t *g = (uint64_t *)(O.events + O.ev_off[sim]);
    for (uint32_t i = (uint32_t)lane; i < c * kRecWords; i += kWave) {
        const uint64_t rec = base + i / kRecWords;
        if (rec < cap) __builtin_nontemporal_store(st->words[w][i], &g[base * kRecWords + i]);
    }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) st->count[w] = 0;
    __builtin_amdgcn_wave_barrier();
}

// Per-simulation statistics a batch lane has gathered since it moved to its
// current simulation.
__device__ __forceinline__ void flush_sim_stats(const DevOut &O, uint32_t sim, uint64_t steps,
                                                uint32_t particles, uint32_t nan, uint32_t capped) {
    unsigned long long *g = (unsigned long long *)O.g_stats + (size_t)sim * ST_COUNT_;
    if (steps) atomicAdd(&g[ST_STEPS], (unsigned long long)steps);
    if (particles) atomicAdd(&g[ST_PARTICLES], (unsigned long long)particles);
    if (nan) atomicAdd(&g[ST_NAN], (unsigned long long)nan);
    if (capped) atomicAdd(&g[ST_CAP], (unsigned long long)capped);
}

// The lanes of one wave integrate particles until the id space is exhausted.
//
// kBatch = false: one simulation, parameters read wave-uniformly from sims[0];
//   per-workgroup statistics go to `stat` (LDS).
// kBatch = true: n_sims simulations of one model concatenated into a virtual
//   id space (offsets `off`, in LDS: simulation s owns virtual ids
//   [off[s], off[s+1])).  A lane that takes a virtual id outside its current
//   simulation's range looks the simulation up, flushes its statistics for
//   the old one and loads the new one's step parameters, Philox key and step
//   cap into registers.  A wave only ever moves forward in the id space, so a
//   lane's simulation index never decreases.
template <class Model, bool kBatch, class St>
__device__ void run_lanes(const DevParams *__restrict__ sims, uint32_t n_sims, const uint64_t *off,
                          uint64_t *stat, St *stage, const BmTables &bm, const DevOut &O) {
    constexpr int NPS = Model::kNormals;   // normals per step; a cycle uses NPS Philox blocks
    unsigned long long *queue = (unsigned long long *)O.queue;
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const uint64_t lane_lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const DevParams &P0 = sims[0];
    const bool resume = !kBatch && O.resume != nullptr;
    const uint64_t N = kBatch ? off[n_sims] : (resume ? O.resume_n : P0.n_particles);
    const uint64_t grab = O.chunk ? (uint64_t)O.chunk : kChunk;   // ids per queue grab (wave-uniform)

    // Per-simulation values (wave-uniform for plain runs, per lane in batches).
    uint32_t k0 = (uint32_t)P0.seed, k1 = (uint32_t)(P0.seed >> 32);
    uint32_t cap = (uint32_t)P0.max_steps;
    uint64_t first = P0.first_particle;
    uint32_t sim = 0;
    uint64_t sim_lo = 0, sim_hi = 0;
    StepParams lp;
    if constexpr (kBatch) lp = StepParams{};
    // Plain runs of models with Model::kVecParams hold the step parameters in
    // vector registers (Model::to_vgprs), not scalar ones: see the models.
    constexpr bool kVecParams = kBatch || Model::kVecParams;
    if constexpr (!kBatch && Model::kVecParams) {
        lp = static_cast<const StepParams &>(P0);
        Model::to_vgprs(lp);
    }

    Model m;
    uint64_t pid = 0;
    uint32_t steps = 0, blk = 0;
    bool active = false, has_particle = false;
    uint64_t q_next = 0, q_end = 0;
    bool exhausted = false;
    uint64_t my_steps = 0;
    uint32_t my_particles = 0, my_nan = 0, my_cap = 0, cycles = 0, my_max = 0;

    auto emit = [&](double t_inj, double t_exit, double r, double theta, double sumac) {
        stage_event(stage, O, pid, sim, t_inj, t_exit, r, theta, sumac);
    };
    auto do_step = [&](const double *z) -> bool {
        if constexpr (kVecParams) return m.step(lp, z, bm.trig, emit);
        else return m.step(P0, z, bm.trig, emit);
    };

    for (;;) {
        // ---- refill idle lanes (wave-uniform control flow)
        const uint64_t idle = __ballot(!active);
        if (idle && !exhausted) {
            const int n_idle = __popcll(idle);
            const int rank = __popcll(idle & lane_lt);
            const uint64_t avail = q_end - q_next;
            uint64_t take = ~0ull;
            if (!active && (uint64_t)rank < avail) take = q_next + rank;
            if (avail < (uint64_t)n_idle) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(queue, (unsigned long long)grab);
                base = __shfl(base, 0);
                if (base >= N) {
                    exhausted = true;
                    q_next = q_end = N;
                } else {
                    const uint64_t end = (base + grab < N) ? base + grab : N;
                    const uint64_t need = (uint64_t)n_idle - avail;
                    if (!active && (uint64_t)rank >= avail) {
                        const uint64_t cand = base + ((uint64_t)rank - avail);
                        if (cand < end) take = cand;
                    }
                    q_next = (base + need < end) ? base + need : end;
                    q_end = end;
                }
            } else {
                q_next += (uint64_t)n_idle;
            }
            if (take != ~0ull) {
                // the lane's previous particle ended: its steps, and whether
                // it stopped on a NaN state
                if (has_particle) {
                    my_steps += steps;
                    my_max = steps > my_max ? steps : my_max;
                    if (isnan(m.r)) ++my_nan;
                }
                has_particle = true;
                steps = 0;
                blk = 0;
                double T;
                if constexpr (kBatch) {
                    if (take >= sim_hi) {
                        if (my_particles) flush_sim_stats(O, sim, my_steps, my_particles, my_nan, my_cap);
                        my_steps = 0;
                        my_particles = my_nan = my_cap = 0;
                        // largest s in [sim, n_sims) with off[s] <= take; off[s+1] > take follows
                        uint32_t lo = sim, hi = n_sims;
                        while (hi - lo > 1) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (off[mid] <= take) lo = mid; else hi = mid;
                        }
                        sim = lo;
                        sim_lo = off[lo];
                        sim_hi = off[lo + 1];
                        const DevParams &Q = sims[sim];
                        lp = static_cast<const StepParams &>(Q);
                        k0 = (uint32_t)Q.seed;
                        k1 = (uint32_t)(Q.seed >> 32);
                        cap = (uint32_t)Q.max_steps;
                        first = Q.first_particle;
                    }
                    pid = first + (O.order ? (uint64_t)O.order[take] : take - sim_lo);
                    const DevParams &Q = sims[sim];
                    T = injection_energy(Q, pid);
                    ++my_particles;
                    active = m.init(lp, T);
                } else if (resume) {
                    // a particle handed on by the previous launch
                    const SpillRec &rec = O.resume[take];
                    m.load(rec.s);
                    pid = rec.pid;
                    steps = rec.steps;
                    blk = rec.blk;
                    active = true;
                } else {
                    pid = first + (O.order ? (uint64_t)O.order[take] : take);
                    T = injection_energy(P0, pid);
                    ++my_particles;
                    active = m.init(P0, T);
                }
            }
        }
        if constexpr (!kBatch) {
            // tail compaction: a sparse wave hands its live particles on and exits
            if (O.spill && exhausted && cycles >= kSpillMinCycles) {
                const uint64_t live = __ballot(active);
                const int n_live = __popcll(live);
                if (n_live > 0 && n_live <= kSpillLanes) {
                    unsigned long long base = 0;
                    if (lane == 0) base = atomicAdd((unsigned long long *)O.spill_count, (unsigned long long)n_live);
                    base = __shfl(base, 0);
                    const uint64_t slot = base + (uint64_t)__popcll(live & lane_lt);
                    if (active && slot < O.spill_cap) {
                        SpillRec &rec = O.spill[slot];
                        m.save(rec.s);
                        rec.pid = pid;
                        rec.steps = steps;
                        rec.blk = blk;
                        active = false;
                        has_particle = false;   // its steps are counted where it ends
                    }
                }
            }
        }
        stage_flush<kBatch>(stage, O, kStageFlush);
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        ++cycles;
        // ---- one cycle: kCycleSteps steps, NPS Philox blocks
        if constexpr (NPS == 1) {
            // 1D: the four normals of the cycle come from one block; every
            // normal's position is known at compile time.
            double z[4];
            if constexpr (!kBatch && !Model::kLatencyBound) {
                // wave-uniform Philox key: an empty asm (it emits nothing)
                // makes the round keys be formed with scalar adds in every
                // cycle rather than kept in 20 scalar registers over the loop
                // (spilled to vector-register lanes, each read a VALU
                // instruction).  A latency-bound model keeps them.
                asm volatile("" : "+s"(k0), "+s"(k1));
            }
            normal_block(pid, blk, k0, k1, bm, z);
            if (__ballot(active && steps + kCycleSteps > cap) == 0) {
                // no lane can reach its step cap in this cycle
#pragma unroll
                for (int s = 0; s < kCycleSteps; ++s) {
                    if (active) {
                        ++steps;
                        if (!do_step(&z[s])) active = false;
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < kCycleSteps; ++s) {
                    if (active) {
                        if (steps >= cap) {
                            ++my_cap;
                            active = false;
                        } else {
                            ++steps;
                            if (!do_step(&z[s])) active = false;
                        }
                    }
                }
            }
        } else if constexpr (NPS == 2) {
            // 2D SolarProp: one block serves two steps (normals 4k .. 4k+3 of
            // the cycle); the loop over step pairs stays rolled.
#pragma unroll 1
            for (int s = 0; s < kCycleSteps; s += 2) {
                double t[4];
                if constexpr (!kBatch) asm volatile("" : "+s"(k0), "+s"(k1));   // as above
                normal_block(pid, blk + (uint32_t)(s >> 1), k0, k1, bm, t);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (active) {
                        if (steps >= cap) {
                            ++my_cap;
                            active = false;
                        } else {
                            ++steps;
                            if (!do_step(&t[2 * h])) active = false;
                        }
                    }
                }
            }
        } else {
            // 2D Geliosphere: steps stay rolled (the step body is large).
            // Step s of the cycle takes normals 3s .. 3s+2: the carried tail
            // of the previous block and the head of the next.  The split
            // depends only on s (wave-uniform), so each phase is a scalar
            // branch with its normals in fixed registers (no selects).
            static_assert(NPS == 3 && kCycleSteps == 4, "three blocks serve the four steps of a cycle");
            double c0 = 0.0, c1 = 0.0, c2 = 0.0;   // unused normals of the last block
#pragma unroll 1
            for (int s = 0; s < kCycleSteps; ++s) {
                double z[3];
                if (s == 3) {
                    z[0] = c0; z[1] = c1; z[2] = c2;
                } else {
                    double t[4];
                    if constexpr (!kBatch) asm volatile("" : "+s"(k0), "+s"(k1));   // as above
                    normal_block(pid, blk + (uint32_t)s, k0, k1, bm, t);
                    if (s == 0) {
                        z[0] = t[0]; z[1] = t[1]; z[2] = t[2];
                        c0 = t[3];
                    } else if (s == 1) {
                        z[0] = c0; z[1] = t[0]; z[2] = t[1];
                        c0 = t[2]; c1 = t[3];
                    } else {
                        z[0] = c0; z[1] = c1; z[2] = t[0];
                        c0 = t[1]; c1 = t[2]; c2 = t[3];
                    }
                }
                if (active) {
                    if (steps >= cap) {
                        ++my_cap;
                        active = false;
                    } else {
                        ++steps;
                        if (!do_step(z)) active = false;
                    }
                }
            }
        }
        blk += (uint32_t)NPS;
    }
    if (has_particle) {
        my_steps += steps;
        my_max = steps > my_max ? steps : my_max;
        if (isnan(m.r)) ++my_nan;
    }
    stage_flush<kBatch>(stage, O, 1);
    if constexpr (kBatch) {
        if (my_particles) flush_sim_stats(O, sim, my_steps, my_particles, my_nan, my_cap);
        if (lane == 0) atomicAdd((unsigned long long *)O.cycles, (unsigned long long)cycles);
    } else {
        atomicAdd((unsigned long long *)&stat[ST_STEPS], (unsigned long long)my_steps);
        atomicAdd((unsigned long long *)&stat[ST_PARTICLES], (unsigned long long)my_particles);
        if (my_nan) atomicAdd((unsigned long long *)&stat[ST_NAN], (unsigned long long)my_nan);
        if (my_cap) atomicAdd((unsigned long long *)&stat[ST_CAP], (unsigned long long)my_cap);
        if (lane == 0) atomicAdd((unsigned long long *)&stat[ST_CYCLES], (unsigned long long)cycles);
        if (my_max) atomicMax((unsigned long long *)&stat[ST_MAX_STEPS], (unsigned long long)my_max);
    }
}

// Flush the workgroup histogram into the global one.  Digits are carry-
// normalised to 32 bits first, so every global digit receives values below
// 2^32 and cannot overflow before 2^32 workgroups have flushed.
__device__ void flush_hist(const DevParams &P, uint64_t *lds, const DevOut &O, uint32_t sim) {
    const int nb = (int)P.n_bins, nw = (int)P.n_weights;
    uint64_t *cnt = lds;
    uint64_t *dig = lds + nb;
    uint64_t *stat = dig + nw * 2 * nb * GS_ACC_LIMBS;
    unsigned long long *g_cnt = (unsigned long long *)O.g_count + (size_t)sim * GS_MAX_BINS;
    unsigned long long *g_dig = (unsigned long long *)O.g_digits + (size_t)sim * kDigitsPerSim;
    unsigned long long *g_stat = (unsigned long long *)O.g_stats + (size_t)sim * ST_COUNT_;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (cnt[i]) atomicAdd(&g_cnt[i], (unsigned long long)cnt[i]);
    for (int a = threadIdx.x; a < nw * 2 * nb; a += blockDim.x) {
        const int k = a / (2 * nb), rest = a % (2 * nb), sh = rest / nb, b = rest % nb;
        const uint64_t *d = dig + (size_t)a * GS_ACC_LIMBS;
        unsigned long long *g = g_dig + ((size_t)(k * 2 + sh) * GS_MAX_BINS + b) * GS_ACC_LIMBS;
        uint64_t carry = 0;
        for (int i = 0; i < GS_ACC_LIMBS; ++i) {
            const uint64_t v = d[i] + carry;
            carry = v >> 32;
            const uint64_t lo = v & 0xffffffffull;
            if (lo) atomicAdd(&g[i], (unsigned long long)lo);
        }
        if (carry) atomicAdd(&g_stat[ST_OVERFLOW], 1ull);
    }
    for (int i = threadIdx.x; i < ST_COUNT_; i += blockDim.x)
        if (stat[i]) atomicAdd(&g_stat[i], (unsigned long long)stat[i]);
}

__device__ __forceinline__ void zero_lds(uint64_t *lds, int words) {
    for (int i = threadIdx.x; i < words; i += blockDim.x) lds[i] = 0;
}

// kMinWaves: waves per SIMD asked of the register allocator (1: no request);
// kThreads: the workgroup size (128, 256 or 512).
template <class Model, int kMinWaves, int kThreads>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kMinWaves)))
plain_kernel(const DevParams *__restrict__ Pp, const DevOut O) {
    __shared__ uint64_t stat[ST_COUNT_];
    __shared__ Stage<kThreads / kWave> stage;
    __shared__ BmTables bm;
    if (threadIdx.x < ST_COUNT_) stat[threadIdx.x] = 0;
    if (threadIdx.x < kThreads / kWave) stage.count[threadIdx.x] = 0;
    load_bm_tables(&bm);
    __syncthreads();
    run_lanes<Model, false>(Pp, 1, nullptr, stat, &stage, bm, O);
    __syncthreads();
    if (threadIdx.x < ST_COUNT_ && stat[threadIdx.x]) {
        unsigned long long *g = (unsigned long long *)&O.g_stats[threadIdx.x];
        if (threadIdx.x == ST_MAX_STEPS) atomicMax(g, (unsigned long long)stat[threadIdx.x]);
        else atomicAdd(g, (unsigned long long)stat[threadIdx.x]);
    }
}

// Batch of n_sims simulations of one model (n_sims <= kBatchMaxSims).
// Lanes hold their simulation's step parameters in VGPRs; kMinWaves asks the
// register allocator for that many waves per SIMD (2D models: 2, which holds
// those parameters without spills).
template <class Model, int kMinWaves, int kThreads>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kMinWaves)))
batch_kernel(const DevParams *__restrict__ sims, uint32_t n_sims, const uint64_t *__restrict__ offsets,
             const DevOut O) {
    __shared__ uint64_t off[kBatchMaxSims + 1];
    __shared__ Stage<kThreads / kWave> stage;
    __shared__ BmTables bm;
    for (uint32_t i = threadIdx.x; i <= n_sims; i += blockDim.x) off[i] = offsets[i];
    if (threadIdx.x < kThreads / kWave) stage.count[threadIdx.x] = 0;
    load_bm_tables(&bm);
    __syncthreads();
    run_lanes<Model, true>(sims, n_sims, off, nullptr, &stage, bm, O);
}

// Histogram pass: workgroup (x, s) folds a strided share of simulation s's
// events into its LDS histogram and flushes it.  Sums are exact, so the order
// of events in a region does not matter.
__global__ void __launch_bounds__(kBlockThreads) fold_kernel(const DevParams *__restrict__ sims,
                                                             const DevOut O) {
    extern __shared__ uint64_t lds[];
    const uint32_t sim = blockIdx.y;
    const DevParams &P = sims[sim];
    uint64_t n = O.n_events[sim];
    const uint64_t cap = O.ev_cap[sim];
    if (n > cap) n = cap;
    const DevEvent *ev = O.events + O.ev_off[sim];
    zero_lds(lds, lds_words((int)P.n_bins, (int)P.n_weights));
    __syncthreads();
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        fold_one(P, lds, ev[i]);
    __syncthreads();
    flush_hist(P, lds, O, sim);
}

// Diagnostics: evaluate one math-layer function elementwise, or dump the
// normals of one particle.
__global__ void debug_kernel(int fn, const double *x, double *y, uint64_t n, uint64_t seed,
                             uint64_t pid) {
    __shared__ BmTables bm;
    load_bm_tables(&bm);
    __syncthreads();
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s, c, z[4];
    switch (fn) {
    case GS_DBG_RCP: y[i] = fm::rcp(x[i]); break;
    case GS_DBG_RSQRT: y[i] = fm::rsqrt(x[i]); break;
    case GS_DBG_SQRT: y[i] = fm::sqrt_pos(x[i]); break;
    case GS_DBG_LOG: y[i] = fm::log_fast(x[i]); break;
    case GS_DBG_EXP: y[i] = fm::exp_fast(x[i]); break;
    case GS_DBG_SIN_TURN: fm::sincos_turnd((uint32_t)x[i], bm.trig, s, c); y[i] = s; break;
    case GS_DBG_COS_TURN: fm::sincos_turnd((uint32_t)x[i], bm.trig, s, c); y[i] = c; break;
    case GS_DBG_LOG_U32: y[i] = fm::log_u32d((uint32_t)x[i], bm.log); break;
    case GS_DBG_RAW_RCP: y[i] = __builtin_amdgcn_rcp(x[i]); break;
    case GS_DBG_RAW_RSQ: y[i] = __builtin_amdgcn_rsq(x[i]); break;
    case GS_DBG_RAW_SQRT: y[i] = __builtin_amdgcn_sqrt(x[i]); break;
    case GS_DBG_SIN: fm::sincosd(x[i], bm.trig, s, c); y[i] = s; break;
    case GS_DBG_COS: fm::sincosd(x[i], bm.trig, s, c); y[i] = c; break;
    case GS_DBG_ATAN: y[i] = fm::atand(x[i]); break;
    case GS_DBG_ATAN_FRAC: y[i] = fm::atan_fracd(x[i], x[n + i], bm.trig + fm::kBmTrigWords); break;
    case GS_DBG_SHEET_F: fm::sheet_drift<fm::DevOps>(x[i], x[n + i], bm.trig + fm::kBmTrigWords, s, c); y[i] = s; break;
    case GS_DBG_SHEET_G: fm::sheet_drift<fm::DevOps>(x[i], x[n + i], bm.trig + fm::kBmTrigWords, s, c); y[i] = c; break;
    case GS_DBG_RSQRT4: y[i] = fm::rsqrt4(x[i]); break;
    case GS_DBG_NORMALS:
        normal_block(pid, (uint32_t)(i >> 2), (uint32_t)seed, (uint32_t)(seed >> 32), bm, z);
        y[i] = z[i & 3];
        break;
    default: y[i] = 0.0; break;
    }
}

// ---------------------------------------------------------------- order
// A plain run's particles sorted by injection energy.  On the backward paths
// low-energy particles diffuse slowest and live longest (two orders of
// magnitude across 0.01-100 GeV), so handing them out first keeps the run
// from ending on a few long trajectories in otherwise idle waves.  Results do
// not depend on the order: particles are keyed by id and sums are exact.
__device__ __forceinline__ uint32_t order_bucket(const DevParams &P, uint64_t i, const OrderKey &key) {
    const double T = injection_energy(P, P.first_particle + i);
    const uint64_t k = (uint64_t)__double_as_longlong(T) >> key.shift;
    const uint64_t b = k > key.kmin ? k - key.kmin : 0;
    return b < key.nb ? (uint32_t)b : key.nb - 1;
}

constexpr int kOrderPerThread = 16;
constexpr uint64_t kOrderTile = (uint64_t)kBlockThreads * kOrderPerThread;

// counts[b] = particles in bucket b.
__global__ void __launch_bounds__(kBlockThreads) order_count_kernel(const DevParams *__restrict__ sims, uint64_t n,
                                                                   OrderKey key, uint32_t *counts) {
    __shared__ uint32_t h[kOrderMaxBuckets];
    for (uint32_t b = threadIdx.x; b < key.nb; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const DevParams &P = sims[0];
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[order_bucket(P, i, key)], 1u);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < key.nb; b += blockDim.x)
        if (h[b]) atomicAdd(&counts[b], h[b]);
}

// cursor[b] = sum of counts[0 .. b) (one workgroup; nb <= kOrderMaxBuckets).
__global__ void __launch_bounds__(kOrderMaxBuckets) order_scan_kernel(const uint32_t *counts, uint32_t *cursor,
                                                                     uint32_t nb) {
    __shared__ uint32_t v[kOrderMaxBuckets];
    const uint32_t t = threadIdx.x;
    v[t] = t < nb ? counts[t] : 0u;
    __syncthreads();
    for (uint32_t d = 1; d < kOrderMaxBuckets; d <<= 1) {
        const uint32_t add = t >= d ? v[t - d] : 0u;
        __syncthreads();
        v[t] += add;
        __syncthreads();
    }
    if (t < nb) cursor[t] = v[t] - counts[t];
}

// Workgroup x places the offsets of tile [x kOrderTile, (x+1) kOrderTile):
// ranks within the tile from LDS counters, one range per bucket reserved
// from the global cursors.
__global__ void __launch_bounds__(kBlockThreads) order_scatter_kernel(const DevParams *__restrict__ sims, uint64_t n,
                                                                     OrderKey key, uint32_t *cursor,
                                                                     uint32_t *perm) {
    __shared__ uint32_t h[kOrderMaxBuckets];
    for (uint32_t b = threadIdx.x; b < key.nb; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const DevParams &P = sims[0];
    const uint64_t base = blockIdx.x * kOrderTile;
    uint32_t bk[kOrderPerThread], rk[kOrderPerThread];
#pragma unroll
    for (int j = 0; j < kOrderPerThread; ++j) {
        const uint64_t i = base + (uint64_t)j * blockDim.x + threadIdx.x;
        bk[j] = 0;
        rk[j] = 0;
        if (i < n) {
            bk[j] = order_bucket(P, i, key);
            rk[j] = atomicAdd(&h[bk[j]], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < key.nb; b += blockDim.x)
        h[b] = h[b] ? atomicAdd(&cursor[b], h[b]) : 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kOrderPerThread; ++j) {
        const uint64_t i = base + (uint64_t)j * blockDim.x + threadIdx.x;
        if (i < n) perm[h[bk[j]] + rk[j]] = (uint32_t)i;
    }
}

// Batch launches: the same order per simulation over the virtual id space.
// Counters per simulation and bucket are global (kOrderSimBuckets per
// simulation); the offsets sit in LDS as in batch_kernel.
__device__ __forceinline__ uint32_t sim_of(const uint64_t *off, uint32_t n_sims, uint64_t v) {
    uint32_t lo = 0, hi = n_sims;   // largest s with off[s] <= v
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kBlockThreads) batch_order_count_kernel(const DevParams *__restrict__ sims,
                                                                         uint32_t n_sims,
                                                                         const uint64_t *__restrict__ offsets,
                                                                         uint32_t *counts) {
    __shared__ uint64_t off[kBatchMaxSims + 1];
    for (uint32_t i = threadIdx.x; i <= n_sims; i += blockDim.x) off[i] = offsets[i];
    __syncthreads();
    const uint64_t n = off[n_sims];
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t sm = sim_of(off, n_sims, v);
        const DevParams &P = sims[sm];
        if (P.order.nb == 0) continue;
        atomicAdd(&counts[(size_t)sm * kOrderSimBuckets + order_bucket(P, v - off[sm], P.order)], 1u);
    }
}

// Workgroup s: exclusive scan of simulation s's bucket counts into cursor.
__global__ void __launch_bounds__(kOrderSimBuckets) batch_order_scan_kernel(const DevParams *__restrict__ sims,
                                                                          const uint32_t *counts, uint32_t *cursor) {
    __shared__ uint32_t v[kOrderSimBuckets];
    const uint32_t sm = blockIdx.x, t = threadIdx.x;
    const uint32_t nb = sims[sm].order.nb;
    const uint32_t *c = counts + (size_t)sm * kOrderSimBuckets;
    v[t] = t < nb ? c[t] : 0u;
    __syncthreads();
    for (uint32_t d = 1; d < kOrderSimBuckets; d <<= 1) {
        const uint32_t add = t >= d ? v[t - d] : 0u;
        __syncthreads();
        v[t] += add;
        __syncthreads();
    }
    if (t < nb) cursor[(size_t)sm * kOrderSimBuckets + t] = v[t] - c[t];
}

__global__ void __launch_bounds__(kBlockThreads) batch_order_scatter_kernel(const DevParams *__restrict__ sims,
                                                                           uint32_t n_sims,
                                                                           const uint64_t *__restrict__ offsets,
                                                                           uint32_t *cursor, uint32_t *perm) {
    __shared__ uint64_t off[kBatchMaxSims + 1];
    for (uint32_t i = threadIdx.x; i <= n_sims; i += blockDim.x) off[i] = offsets[i];
    __syncthreads();
    const uint64_t n = off[n_sims];
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t sm = sim_of(off, n_sims, v);
        const DevParams &P = sims[sm];
        const uint64_t local = v - off[sm];
        uint64_t pos = local;
        if (P.order.nb != 0)
            pos = atomicAdd(&cursor[(size_t)sm * kOrderSimBuckets + order_bucket(P, local, P.order)], 1u);
        perm[off[sm] + pos] = (uint32_t)local;
    }
}

__global__ void zero_kernel(uint64_t *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace gs

// ---------------------------------------------------------------- scratch
// The no-allocation plain run's working memory: a fixed array in every
// device's global memory, held until the process ends.  Host code must not
// use its name as an address: the symbol's device address is looked up once
// per device (hipGetSymbolAddress, with that device current), and that
// address is what kernels and copies receive.
static __device__ __align__(256) uint64_t scratch[gs::kScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t gs_scratch_address(int device, void **ptr, size_t *bytes) {
    if (device < 0 || device >= kMaxDevices || !ptr || !bytes) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    *bytes = sizeof(scratch);
    return hipSuccess;
}

// ---------------------------------------------------------------- launchers
using namespace gs;

// Waves per SIMD asked of the register allocator, per kernel (1: no request;
// the 1D kernels reach 4 waves unasked, SolarProp 3).  Measured: SolarProp at
// 4 waves (128 VGPRs) spills and runs 5 % slower.
#ifndef GS_BATCH_WAVES_GE
#define GS_BATCH_WAVES_GE 2   // 227 VGPRs, no spills; at 3 waves 62 spill and the batch ran 12 % slower
#endif
#ifndef GS_PLAIN_WAVES_1D
#define GS_PLAIN_WAVES_1D 1
#endif
#ifndef GS_PLAIN_WAVES_BP
#define GS_PLAIN_WAVES_BP 5   // 96 VGPRs, no spills: +3 % over the 98-VGPR, 4-wave shape
#endif
#ifndef GS_PLAIN_WAVES_GE
#define GS_PLAIN_WAVES_GE 3   // 154 VGPRs, no spills
#endif
#ifndef GS_PLAIN_WAVES_SP
#define GS_PLAIN_WAVES_SP 1
#endif
#ifndef GS_BATCH_WAVES_BP
#define GS_BATCH_WAVES_BP 5   // 96 VGPRs (17 spilled); 4 waves (121, none) ran 0.8 % faster on 32 B-p simulations at 30 % of the batch benchmark's counts, not switched: the tuned batch shapes were measured at 5
#endif
constexpr int kBatchWaves1D = 1, kBatchWavesBp = GS_BATCH_WAVES_BP, kBatchWavesSp = 2, kBatchWavesGe = GS_BATCH_WAVES_GE;
constexpr int kPlainWavesBp = GS_PLAIN_WAVES_BP;
constexpr int kPlainWaves1D = GS_PLAIN_WAVES_1D, kPlainWavesSp = GS_PLAIN_WAVES_SP, kPlainWavesGe = GS_PLAIN_WAVES_GE;

// A batch workgroup holds 8 KB of simulation offsets in LDS: at 128 threads
// the LDS allows 4 waves per SIMD at most, so no more are asked for.
constexpr int batch_waves(int waves, int threads) { return threads == 128 && waves > 4 ? 4 : waves; }

template <class Model, int kBatchWaves, int kPlainWaves, int kThreads>
static const void *traj_kernel_ptr_t(bool batch) {
    return batch ? (const void *)&batch_kernel<Model, batch_waves(kBatchWaves, kThreads), kThreads>
                 : (const void *)&plain_kernel<Model, kPlainWaves, kThreads>;
}

template <int kThreads>
static const void *traj_kernel_ptr_b(int model, bool batch) {
    switch (model) {
    case GS_MODEL_1D_FP: return traj_kernel_ptr_t<Fp, kBatchWaves1D, kPlainWaves1D, kThreads>(batch);
    case GS_MODEL_1D_BP: return traj_kernel_ptr_t<Bp, kBatchWavesBp, kPlainWavesBp, kThreads>(batch);
    case GS_MODEL_2D_SOLARPROP: return traj_kernel_ptr_t<Sp, kBatchWavesSp, kPlainWavesSp, kThreads>(batch);
    default: return traj_kernel_ptr_t<Ge, kBatchWavesGe, kPlainWavesGe, kThreads>(batch);
    }
}

static const void *traj_kernel_ptr(int model, bool batch, unsigned block_threads) {
    switch (block_threads) {
    case 128: return traj_kernel_ptr_b<128>(model, batch);
    case 512: return traj_kernel_ptr_b<512>(model, batch);
    default: return traj_kernel_ptr_b<256>(model, batch);
    }
}

bool gs_traj_block_supported(unsigned block_threads) {
    return block_threads == 128 || block_threads == 256 || block_threads == 512;
}

hipError_t gs_launch_debug(int fn, const double *x, double *y, uint64_t n, uint64_t seed,
                           uint64_t pid, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(debug_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fn, x, y, n,
                       seed, pid);
    return hipGetLastError();
}

hipError_t gs_launch_zero(uint64_t *p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t gs_traj_occupancy(int model, bool batch, unsigned block_threads, int *blocks_per_cu) {
    if (!gs_traj_block_supported(block_threads)) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, traj_kernel_ptr(model, batch, block_threads),
                                                        (int)block_threads, 0);
}

template <class Model, int kBatchWaves, int kPlainWaves, int kThreads>
static void launch_traj_t(bool batch, const DevParams *sims, uint32_t n_sims, const uint64_t *offsets,
                          const DevOut &O, unsigned grid, hipStream_t s) {
    if (batch)
        hipLaunchKernelGGL((batch_kernel<Model, batch_waves(kBatchWaves, kThreads), kThreads>), dim3(grid),
                           dim3(kThreads), 0, s, sims, n_sims, offsets, O);
    else
        hipLaunchKernelGGL((plain_kernel<Model, kPlainWaves, kThreads>), dim3(grid), dim3(kThreads), 0, s, sims, O);
}

template <int kThreads>
static bool launch_traj_b(int model, bool batch, const DevParams *sims, uint32_t n_sims, const uint64_t *offsets,
                          const DevOut &O, unsigned grid, hipStream_t s) {
    switch (model) {
    case GS_MODEL_1D_FP:
        launch_traj_t<Fp, kBatchWaves1D, kPlainWaves1D, kThreads>(batch, sims, n_sims, offsets, O, grid, s);
        return true;
    case GS_MODEL_1D_BP:
        launch_traj_t<Bp, kBatchWavesBp, kPlainWavesBp, kThreads>(batch, sims, n_sims, offsets, O, grid, s);
        return true;
    case GS_MODEL_2D_SOLARPROP:
        launch_traj_t<Sp, kBatchWavesSp, kPlainWavesSp, kThreads>(batch, sims, n_sims, offsets, O, grid, s);
        return true;
    case GS_MODEL_2D_GELIOSPHERE:
        launch_traj_t<Ge, kBatchWavesGe, kPlainWavesGe, kThreads>(batch, sims, n_sims, offsets, O, grid, s);
        return true;
    default: return false;
    }
}

hipError_t gs_launch_traj(int model, bool batch, const DevParams *sims, uint32_t n_sims,
                          const uint64_t *offsets, const DevOut &O, unsigned grid, unsigned block_threads,
                          hipStream_t s) {
    if (n_sims == 0 || grid == 0) return hipSuccess;
    if ((!batch && n_sims != 1) || (batch && (n_sims > (uint32_t)kBatchMaxSims || !offsets)))
        return hipErrorInvalidValue;
    bool ok;
    switch (block_threads) {
    case 128: ok = launch_traj_b<128>(model, batch, sims, n_sims, offsets, O, grid, s); break;
    case 256: ok = launch_traj_b<256>(model, batch, sims, n_sims, offsets, O, grid, s); break;
    case 512: ok = launch_traj_b<512>(model, batch, sims, n_sims, offsets, O, grid, s); break;
    default: ok = false; break;
    }
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t gs_launch_order(const DevParams *sims, uint64_t n, const OrderKey &key, uint32_t *counts,
                           uint32_t *perm, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (key.nb == 0 || key.nb > kOrderMaxBuckets || n > 0xffffffffull || ((uintptr_t)counts & 7u))
        return hipErrorInvalidValue;
    uint32_t *cursor = counts + key.nb;
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(kBlockThreads), 0, s, (uint64_t *)counts, (size_t)key.nb);
    const uint64_t blocks = (n + kBlockThreads - 1) / kBlockThreads;
    hipLaunchKernelGGL(order_count_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(kBlockThreads), 0, s,
                       sims, n, key, counts);
    hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(kOrderMaxBuckets), 0, s, counts, cursor, key.nb);
    hipLaunchKernelGGL(order_scatter_kernel, dim3((unsigned)((n + kOrderTile - 1) / kOrderTile)), dim3(kBlockThreads),
                       0, s, sims, n, key, cursor, perm);
    return hipGetLastError();
}

hipError_t gs_launch_batch_order(const DevParams *sims, uint32_t n_sims, const uint64_t *offsets, uint64_t n,
                                 uint32_t *counts, uint32_t *perm, hipStream_t s) {
    if (n_sims == 0 || n == 0) return hipSuccess;
    if (n_sims > (uint32_t)kBatchMaxSims || ((uintptr_t)counts & 7u)) return hipErrorInvalidValue;
    uint32_t *cursor = counts + (size_t)n_sims * kOrderSimBuckets;
    const uint64_t words = (uint64_t)n_sims * kOrderSimBuckets / 2;   // counts, as 64-bit words
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<uint64_t>((words + 255) / 256, 1024)), dim3(256), 0, s,
                       (uint64_t *)counts, (size_t)words);
    const uint64_t blocks = std::min<uint64_t>((n + kBlockThreads - 1) / kBlockThreads, 8192);
    hipLaunchKernelGGL(batch_order_count_kernel, dim3((unsigned)blocks), dim3(kBlockThreads), 0, s, sims, n_sims,
                       offsets, counts);
    hipLaunchKernelGGL(batch_order_scan_kernel, dim3(n_sims), dim3(kOrderSimBuckets), 0, s, sims, counts, cursor);
    hipLaunchKernelGGL(batch_order_scatter_kernel, dim3((unsigned)blocks), dim3(kBlockThreads), 0, s, sims, n_sims,
                       offsets, cursor, perm);
    return hipGetLastError();
}

hipError_t gs_launch_fold(const DevParams *sims, uint32_t n_sims, const DevOut &O, size_t lds_bytes,
                          unsigned blocks_per_sim, hipStream_t s) {
    if (n_sims == 0 || blocks_per_sim == 0) return hipSuccess;
    if (n_sims > 65535u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fold_kernel, dim3(blocks_per_sim, n_sims), dim3(kBlockThreads), lds_bytes, s,
                       sims, O);
    return hipGetLastError();
}
