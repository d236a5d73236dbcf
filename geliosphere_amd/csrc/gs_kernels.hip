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
#include "gs_device.hpp"
#include "gs_kernels.h"

namespace gs {

// Append one event to the event buffer (counted even when the buffer is full).
__device__ __forceinline__ void emit(const DevOut &O, const Event &ev, uint64_t pid, uint32_t sim) {
    const unsigned long long i = atomicAdd((unsigned long long *)O.n_events, 1ull);
    if (i < O.ev_cap) {
        DevEvent d;
        d.t_inj = ev.t_inj; d.t_exit = ev.t_exit; d.r = ev.r; d.theta = ev.theta;
        d.sumac = ev.sumac; d.id = pid; d.sim = sim; d.pad = 0;
        O.events[i] = d;
    }
}

template <class Model>
__device__ void run_lanes(const DevParams &P, uint64_t *stat, unsigned long long *queue,
                          const DevOut &O, uint32_t sim) {
    constexpr int NPS = Model::kNormals;   // normals per step; a cycle uses NPS Philox blocks
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const uint64_t lane_lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t k0 = (uint32_t)P.seed, k1 = (uint32_t)(P.seed >> 32);
    const uint64_t N = P.n_particles;
    const uint32_t cap = (uint32_t)P.max_steps;

    Model m;
    uint64_t pid = 0;
    uint32_t steps = 0, blk = 0;
    bool active = false;
    uint64_t q_next = 0, q_end = 0;
    bool exhausted = false;
    uint64_t my_steps = 0;
    uint32_t my_particles = 0, my_nan = 0, my_cap = 0, cycles = 0;

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
                if (lane == 0) base = atomicAdd(queue, (unsigned long long)kChunk);
                base = __shfl(base, 0);
                if (base >= N) {
                    exhausted = true;
                    q_next = q_end = N;
                } else {
                    const uint64_t end = (base + kChunk < N) ? base + kChunk : N;
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
                my_steps += steps;
                pid = P.first_particle + take;
                steps = 0;
                blk = 0;
                ++my_particles;
                active = m.init(P, injection_energy(P, pid));
            }
        }
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
            normal_block(pid, blk, k0, k1, z);
#pragma unroll
            for (int s = 0; s < kCycleSteps; ++s) {
                if (active) {
                    if (steps >= cap) {
                        ++my_cap;
                        active = false;
                    } else {
                        Event ev;
                        const int rc = m.step(P, &z[s], ev);
                        ++steps;
                        if (rc == STEP_EVENT || rc == STEP_EVENT_DONE) emit(O, ev, pid, sim);
                        if (rc >= STEP_EVENT_DONE) {
                            active = false;
                            if (rc == STEP_DONE && isnan(m.r)) ++my_nan;
                        }
                    }
                }
            }
        } else {
            // 2D: steps stay rolled (the step body is large); normals come from
            // a four-entry buffer refilled from the next block when empty.  All
            // lanes are block-aligned at the cycle start, so the refill branch
            // is wave-uniform.
            double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
            int used = 4;
            uint32_t nb = blk;
#pragma unroll 1
            for (int s = 0; s < kCycleSteps; ++s) {
                double z[NPS];
#pragma unroll
                for (int i = 0; i < NPS; ++i) {
                    if (used == 4) {
                        double t[4];
                        normal_block(pid, nb, k0, k1, t);
                        b0 = t[0]; b1 = t[1]; b2 = t[2]; b3 = t[3];
                        ++nb;
                        used = 0;
                    }
                    z[i] = used == 0 ? b0 : used == 1 ? b1 : used == 2 ? b2 : b3;
                    ++used;
                }
                if (active) {
                    if (steps >= cap) {
                        ++my_cap;
                        active = false;
                    } else {
                        Event ev;
                        const int rc = m.step(P, z, ev);
                        ++steps;
                        if (rc == STEP_EVENT || rc == STEP_EVENT_DONE) emit(O, ev, pid, sim);
                        if (rc >= STEP_EVENT_DONE) {
                            active = false;
                            if (rc == STEP_DONE && isnan(m.r)) ++my_nan;
                        }
                    }
                }
            }
        }
        blk += (uint32_t)NPS;
    }
    my_steps += steps;
    atomicAdd((unsigned long long *)&stat[ST_STEPS], (unsigned long long)my_steps);
    atomicAdd((unsigned long long *)&stat[ST_PARTICLES], (unsigned long long)my_particles);
    if (my_nan) atomicAdd((unsigned long long *)&stat[ST_NAN], (unsigned long long)my_nan);
    if (my_cap) atomicAdd((unsigned long long *)&stat[ST_CAP], (unsigned long long)my_cap);
    if (lane == 0) atomicAdd((unsigned long long *)&stat[ST_CYCLES], (unsigned long long)cycles);
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

template <class Model>
__global__ void __launch_bounds__(kBlockThreads) plain_kernel(const DevParams *__restrict__ Pp,
                                                              const DevOut O) {
    __shared__ uint64_t stat[ST_COUNT_];
    const DevParams &P = *Pp;
    if (threadIdx.x < ST_COUNT_) stat[threadIdx.x] = 0;
    __syncthreads();
    run_lanes<Model>(P, stat, (unsigned long long *)O.queue, O, 0);
    __syncthreads();
    if (threadIdx.x < ST_COUNT_ && stat[threadIdx.x])
        atomicAdd((unsigned long long *)&O.g_stats[threadIdx.x], (unsigned long long)stat[threadIdx.x]);
}

// Histogram pass: every workgroup folds a strided share of the events into its
// LDS histogram and flushes it.  Sums are exact, so the order of events in the
// buffer does not matter.
__global__ void __launch_bounds__(kBlockThreads) fold_kernel(const DevParams *__restrict__ Pp,
                                                             const DevOut O, uint64_t n) {
    extern __shared__ uint64_t lds[];
    const DevParams &P = *Pp;
    zero_lds(lds, lds_words((int)P.n_bins, (int)P.n_weights));
    __syncthreads();
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        fold_one(P, lds, O.events[i]);
    __syncthreads();
    flush_hist(P, lds, O, 0);
}

// Diagnostics: evaluate one math-layer function elementwise, or dump the
// normals of one particle.
__global__ void debug_kernel(int fn, const double *x, double *y, uint64_t n, uint64_t seed,
                             uint64_t pid) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s, c, z[4];
    switch (fn) {
    case GS_DBG_RCP: y[i] = fm::rcp(x[i]); break;
    case GS_DBG_RSQRT: y[i] = fm::rsqrt(x[i]); break;
    case GS_DBG_SQRT: y[i] = fm::sqrt_pos(x[i]); break;
    case GS_DBG_LOG: y[i] = fm::log_fast(x[i]); break;
    case GS_DBG_EXP: y[i] = fm::exp_fast(x[i]); break;
    case GS_DBG_SIN_TURN: fm::sincos_turnd((uint32_t)x[i], s, c); y[i] = s; break;
    case GS_DBG_COS_TURN: fm::sincos_turnd((uint32_t)x[i], s, c); y[i] = c; break;
    case GS_DBG_LOG_U32: y[i] = fm::log_u32d((uint32_t)x[i]); break;
    case GS_DBG_RAW_RCP: y[i] = __builtin_amdgcn_rcp(x[i]); break;
    case GS_DBG_RAW_RSQ: y[i] = __builtin_amdgcn_rsq(x[i]); break;
    case GS_DBG_RAW_SQRT: y[i] = __builtin_amdgcn_sqrt(x[i]); break;
    case GS_DBG_NORMALS:
        normal_block(pid, (uint32_t)(i >> 2), (uint32_t)seed, (uint32_t)(seed >> 32), z);
        y[i] = z[i & 3];
        break;
    default: y[i] = 0.0; break;
    }
}

__global__ void zero_kernel(uint64_t *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace gs

// ---------------------------------------------------------------- launchers
using namespace gs;

static const void *plain_kernel_ptr(int model) {
    switch (model) {
    case GS_MODEL_1D_FP: return (const void *)&plain_kernel<Fp>;
    case GS_MODEL_1D_BP: return (const void *)&plain_kernel<Bp>;
    case GS_MODEL_2D_SOLARPROP: return (const void *)&plain_kernel<Sp>;
    default: return (const void *)&plain_kernel<Ge>;
    }
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

hipError_t gs_plain_occupancy(int model, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, plain_kernel_ptr(model),
                                                        kBlockThreads, 0);
}

hipError_t gs_launch_fold(const DevParams &P, const DevParams *dP, const DevOut &O, uint64_t n,
                          hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t lds = (size_t)lds_words((int)P.n_bins, (int)P.n_weights) * sizeof(uint64_t);
    uint64_t blocks = (n + kBlockThreads - 1) / kBlockThreads;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(fold_kernel, dim3((unsigned)blocks), dim3(kBlockThreads), lds, s, dP, O, n);
    return hipGetLastError();
}

hipError_t gs_launch_plain(const DevParams &P, const DevParams *dP, const DevOut &O, unsigned grid,
                           hipStream_t s) {
    const size_t lds = 0;
    switch (P.model) {
    case GS_MODEL_1D_FP:
        hipLaunchKernelGGL(plain_kernel<Fp>, dim3(grid), dim3(kBlockThreads), lds, s, dP, O);
        break;
    case GS_MODEL_1D_BP:
        hipLaunchKernelGGL(plain_kernel<Bp>, dim3(grid), dim3(kBlockThreads), lds, s, dP, O);
        break;
    case GS_MODEL_2D_SOLARPROP:
        hipLaunchKernelGGL(plain_kernel<Sp>, dim3(grid), dim3(kBlockThreads), lds, s, dP, O);
        break;
    default:
        hipLaunchKernelGGL(plain_kernel<Ge>, dim3(grid), dim3(kBlockThreads), lds, s, dP, O);
        break;
    }
    return hipGetLastError();
}
