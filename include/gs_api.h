/*
 * gs_api.h - C ABI of the gfx950 SDE trajectory engine (heliospheric
 * cosmic-ray modulation, four models: 1D F-p, 1D B-p, 2D SolarProp-like,
 * 2D Geliosphere).
 *
 * Plain C types only: no HIP, torch or C++ types cross this boundary.
 * Every call returns gs_status; GS_OK == 0.  A gs_ctx is bound to one device
 * and is not re-entrant; distinct contexts may be used from distinct threads.
 *
 * Which reference interface each entry point replaces:
 *
 *   gs_run_plain   replaces XGpuModel::prepareAndRunSimulation(ParamsCarrier*)
 *                  of the four GPU models (reference
 *                  CUDAKernel/include/AbstractGpuSimulation.hpp:31;
 *                  CUDAKernel/src/OneDimensionFpGpuModel.cpp:7-40,
 *                  OneDimensionBpGpuModel.cpp:7-43,
 *                  SolarPropLikeGpuModel.cpp:7-44,
 *                  GeliosphereGpuModel.cpp:7-44) together with the
 *                  run*Simulation launch loops they call
 *                  (OneDimensionBpModel.cu:149-200 etc.) and the log.dat
 *                  re-read + binning of the Results pass
 *                  (Algorithm/src/OneDimensionBpResults.cpp:14-113,
 *                  OneDimensionFpResults.cpp:13-114,
 *                  TwoDimensionBpResults.cpp:13-134): the histogram is built
 *                  on the device, exit records are optional.
 *   gs_params      replaces the __constant__ parameter block filled by
 *                  setConstants (CUDAKernel/src/CosmicConstants.cu:14-111)
 *                  and the simulationInputX structs
 *                  (CUDAKernel/include/OneDimensionBpModel.cuh:69-81).
 *   gs_run_batch   replaces the sequential row loop of
 *                  BatchRun::runAlgorithm (Algorithm/src/BatchRun.cpp:50-78).
 *   gs_create /    replace the per-run cudaMalloc/cudaFree of the GPU model
 *   gs_destroy     wrappers (OneDimensionBpGpuModel.cpp:16-42).
 *   gs_last_error  replaces gpuAssert's print-and-exit
 *                  (CUDAKernel/include/CudaErrorCheck.cuh:25-41).
 */
#ifndef GS_API_H
#define GS_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_ABI_VERSION 1

#define GS_MAX_BINS 64
#define GS_MAX_WEIGHTS 4
/* Exact weight accumulator: GS_ACC_LIMBS limbs of 32 bits; bit 0 of limb 0
 * is worth 2^GS_ACC_LSB_EXP.  Values below 2^GS_ACC_LSB_EXP are truncated
 * (counted in n_acc_inexact), totals must stay below 2^(GS_ACC_LSB_EXP+512). */
#define GS_ACC_LIMBS 16
#define GS_ACC_LSB_EXP (-320)

typedef enum gs_status {
    GS_OK = 0,
    GS_ERR_INVALID_ARG = 1,
    GS_ERR_HIP = 2,
    GS_ERR_OUT_OF_MEMORY = 3,
    GS_ERR_OVERFLOW = 4,
    GS_ERR_NO_DEVICE = 5,
    GS_ERR_IO = 6,
    GS_ERR_UNSUPPORTED = 7
} gs_status;

typedef enum gs_model {
    GS_MODEL_1D_FP = 0,          /* "1D Fp"  forward-in-time  (-F) */
    GS_MODEL_1D_BP = 1,          /* "1D Bp"  backward-in-time (-B) */
    GS_MODEL_2D_SOLARPROP = 2,   /* "2D SolarProp-like"       (-E) */
    GS_MODEL_2D_GELIOSPHERE = 3  /* "2D Geliosphere"          (-T) */
} gs_model;

typedef enum gs_injection {
    /* T = T_min (T_max/T_min)^u, u from the particle's Philox injection draw */
    GS_INJ_LOG_UNIFORM = 0,
    /* T = T_min + (id mod grid_n) (T_max - T_min) / grid_n
     * (reference CpuImplementations/include/Constants.hpp:234-238) */
    GS_INJ_GRID = 1,
    /* T = SolarProp energy table[id mod 30]
     * (reference CpuImplementations/include/Constants.hpp:117-121) */
    GS_INJ_SOLARPROP_BINS = 2,
    /* T = T_min for every particle */
    GS_INJ_FIXED = 3
} gs_injection;

/* Local interstellar spectra used to weight an event.  J is a differential
 * intensity; weight = J_LIS(T_lis) / p_lis^2 * p_obs^2 where (T_lis, p_lis)
 * is the particle at the outer boundary and p_obs at the observer. */
typedef enum gs_weight {
    GS_W_POWERLAW = 0,  /* J = (T + 0.94)^-2.78                              */
    GS_W_YAMADA = 1,    /* f = (m0^2c^4 + p^2c^2)^-1.85 / p  (SI)
                           (Algorithm/src/OneDimensionBpResults.cpp:58-60)   */
    GS_W_WEBER = 2,     /* Webber-Higbie, at T_inj (TwoDimensionBpResults.cpp:50-54) */
    GS_W_USOSKIN = 3,   /* Usoskin 2005 (TwoDimensionBpResults.cpp:59-67)    */
    GS_W_BURGER = 4     /* Burger 2000  (TwoDimensionBpResults.cpp:72-78)    */
} gs_weight;

/* gs_params.flags */
#define GS_FLAG_SOLARPROP_GPU_DRIFT 0x1u /* 2D SolarProp-like: theta drift with
                                            polarity fixed to +1, as the
                                            reference GPU kernel does
                                            (SolarPropLikeModel.cu:213,
                                            CosmicConstants.cu:37) */

typedef struct gs_params {
    int32_t model;               /* gs_model */
    int32_t injection;           /* gs_injection */
    uint32_t flags;
    uint32_t n_bins;             /* 1 .. GS_MAX_BINS */
    uint32_t n_weights;          /* 1 .. GS_MAX_WEIGHTS */
    int32_t weights[GS_MAX_WEIGHTS]; /* gs_weight per weight slot */
    uint32_t grid_n;             /* GS_INJ_GRID only */
    double K0;          /* diffusion coefficient [AU^2/s]                  */
    double V;           /* solar wind speed [AU/s], > 0; 1D B-p uses -V      */
    double dt;          /* time step [s]                                     */
    double r_inj;       /* injection radius [AU] (backward models)           */
    double theta_inj;   /* injection colatitude [rad] (2D models)            */
    double tilt;        /* current-sheet tilt alpha_M [rad] (2D models)      */
    double polarity;    /* +1 / -1 (2D models)                               */
    double ratio;       /* K_perp / K_par (2D models)                        */
    double delta0;      /* polar-field constant C_delta (2D Geliosphere)     */
    double T_min;       /* injection energy range [GeV]                      */
    double T_max;
    uint64_t n_particles;    /* particles of this run                        */
    uint64_t first_particle; /* id of the first particle (shards)            */
    uint64_t seed;           /* Philox key                                   */
    uint64_t max_steps;      /* per-particle step cap, 0 = library default   */
    double bin_edges[GS_MAX_BINS + 1]; /* ascending; bin i = [e_i, e_{i+1}) */
} gs_params;

typedef struct gs_spectrum {
    uint32_t n_bins;
    uint32_t n_weights;
    /* events per bin: exits (B-p, 2D) or 1 AU crossings (F-p) */
    uint64_t count[GS_MAX_BINS];
    /* correctly rounded values of the exact sums below */
    double sum_w[GS_MAX_WEIGHTS][GS_MAX_BINS];
    double sum_w2[GS_MAX_WEIGHTS][GS_MAX_BINS];
    uint64_t n_particles;    /* particles integrated                         */
    uint64_t n_events;       /* events recorded (all bins + outside)         */
    uint64_t n_outside;      /* events whose bin variable is outside edges   */
    uint64_t total_steps;    /* Euler-Maruyama steps over all particles      */
    uint64_t n_nan;          /* particles stopped by a NaN state             */
    uint64_t n_step_cap;     /* particles stopped by max_steps               */
    uint64_t n_acc_inexact;  /* weights truncated below the accumulator LSB  */
    uint64_t n_acc_overflow; /* accumulator overflowed (sums invalid)        */
    /* exact sums, normalised 32-bit limbs, little-endian; [w][0]=sum w,
     * [w][1]=sum w^2 */
    uint32_t acc[GS_MAX_WEIGHTS][2][GS_MAX_BINS][GS_ACC_LIMBS];
} gs_spectrum;

/* One exit (B-p, 2D) or one 1 AU crossing (F-p). 48 bytes. */
typedef struct gs_exit_record {
    double t_inj;    /* injection kinetic energy [GeV]                       */
    double t_exit;   /* kinetic energy at the event [GeV]                    */
    double r;        /* radius at the event [AU]                             */
    double theta;    /* colatitude at the event [rad] (2D), 0 otherwise      */
    double sumac;    /* F-p adiabatic loss integral, 0 otherwise             */
    uint64_t id;     /* particle id                                          */
} gs_exit_record;

typedef struct gs_run_stats {
    double kernel_ms;        /* trajectory kernel time (HIP events)          */
    double total_ms;         /* whole call                                   */
    uint64_t lane_cycles;    /* lane-cycles issued (active + idle)           */
    uint64_t active_cycles;  /* lane-cycles with a live particle             */
    uint32_t grid_blocks;
    uint32_t block_threads;
    uint64_t max_particle_steps; /* plain runs: most steps of one particle     */
} gs_run_stats;

typedef struct gs_ctx gs_ctx;

/* Open a context on HIP device `device`. */
gs_status gs_create(gs_ctx **ctx, int device);
void gs_destroy(gs_ctx *ctx);
/* Message of the last failing call on ctx (never NULL). */
const char *gs_last_error(const gs_ctx *ctx);
int gs_abi_version(void);

/* Fill *p with the defaults of `model` (K0 = 5e22 cm^2/s, V = 400 km/s,
 * dt = 500 s (1D) / 50 s (2D), r = 1 AU, theta = 90 deg, log-uniform
 * 0.01-100 GeV injection, 30 log bins, power-law LIS). Host only. */
gs_status gs_default_params(gs_params *p, int32_t model);
/* Validate parameters without running. Host only. */
gs_status gs_validate_params(const gs_params *p, char *msg, uint32_t msg_len);

/* Run one simulation: particles [first_particle, first_particle+n_particles).
 * out: histogram (required).  recs/rec_cap/n_recs: optional exit records
 * (recs may be NULL); *n_recs receives the number of events, which can
 * exceed rec_cap (only rec_cap are stored, sorted by particle id, then by
 * event order within a particle). */
gs_status gs_run_plain(gs_ctx *ctx, const gs_params *p, gs_spectrum *out,
                       gs_exit_record *recs, uint64_t rec_cap,
                       uint64_t *n_recs);

/* gs_run_plain without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It makes no
 * hipMalloc, hipFree or other allocating call.  Its working memory is a fixed
 * scratch array of 34.2 MiB in each device's global memory, held until the
 * process ends: one simulation's histogram, parameters and counters, the
 * energy order of 2^19 particles and an event region of 2^19 records.  A run
 * goes through it in launches of at most 2^19 particles (1D F-p: 2^23, halved
 * while a launch's events overflow the region).  Histograms up to
 * GS_MAX_BINS x GS_MAX_WEIGHTS fit; larger ones, and a single particle with
 * more than 2^19 events, are rejected (GS_ERR_INVALID_ARG, GS_ERR_OVERFLOW).
 * The result equals gs_run_plain(ctx, p, out, NULL, 0, NULL) bit for bit; no
 * exit records are returned.  Every call on a device shares its scratch,
 * which keeps the last call's contents (each call clears what it uses):
 * calls that overlap on one device, from any contexts or threads, race. */
gs_status gs_run_plain_noalloc(gs_ctx *ctx, const gs_params *p, gs_spectrum *out);

/* Run n_sims independent simulations (any mix of models) together: one
 * persistent launch per model over the concatenated particles of its
 * simulations, the model groups concurrently, then one histogram pass; the
 * launch sequence is a HIP graph reused while the batch layout stays the
 * same.  outs[i] is identical, bit for bit, to gs_run_plain(sims[i]).
 * Simulations whose event region would exceed half the batch event budget
 * run alone.  gs_last_run_stats covers the whole batch. */
gs_status gs_run_batch(gs_ctx *ctx, const gs_params *sims, uint32_t n_sims,
                       gs_spectrum *outs);
/* gs_run_batch that also returns exit records: for simulation i with
 * recs[i] != NULL and rec_caps[i] > 0, recs[i] receives up to rec_caps[i] of
 * its records in gs_run_plain's order (by particle id, then event order);
 * n_recs (may be NULL) receives every simulation's event count.  recs may be
 * NULL (no records).  Replaces what each row of the reference's batch leaves
 * for its Results pass: a full Algorithm run, log.dat included
 * (Algorithm/src/BatchRun.cpp:50-78, OneDimensionBpModel.cu:186-199). */
gs_status gs_run_batch_records(gs_ctx *ctx, const gs_params *sims, uint32_t n_sims, gs_spectrum *outs,
                               gs_exit_record *const *recs, const uint64_t *rec_caps, uint64_t *n_recs);
/* Event records (64 B each) a batch may hold on the device at once
 * (default 2^29); larger batches run as consecutive sub-batches. */
gs_status gs_set_batch_budget(gs_ctx *ctx, uint64_t max_events);

/* Memory footprint of gs_run_plain: particles per trajectory launch
 * (1 .. 2^27; 0 = 2^27) and most exit records held on the device at once
 * (0 = room for every event of a launch).  A launch whose events exceed the
 * limit runs again as half as many particles; results do not change.
 * Replaces the reference's fixed blockSize x threadSize launches and their
 * record buffers of the same size (CUDAKernel/src/OneDimensionFpGpuModel.cpp:19-20,
 * OneDimensionFpModel.cu:172-184). */
gs_status gs_set_plain_limits(gs_ctx *ctx, uint64_t max_chunk, uint64_t max_events);

/* Launch shape of a trajectory kernel.  block_threads: workgroup size (128,
 * 256 or 512); blocks_per_cu: persistent blocks per CU, so that each lane
 * integrates about N / (blocks_per_cu x CUs x block_threads) particles one
 * after another (capped at the kernel's occupancy); chunk: particle ids a
 * wave takes from the work queue per grab.  0 in a field: the default (256,
 * the occupancy, 64).  Results never depend on the shape. */
typedef struct gs_launch_shape {
    uint32_t block_threads;
    uint32_t blocks_per_cu;
    uint32_t chunk;
} gs_launch_shape;
#define GS_KERNEL_PLAIN 0   /* gs_run_plain's trajectory kernel of a model   */
#define GS_KERNEL_BATCH 1   /* gs_run_batch's trajectory kernel of a model   */

/* Override the launch shape used for `model` by both kernels (0 = automatic);
 * block_threads 0, 128, 256 or 512. */
gs_status gs_set_launch(gs_ctx *ctx, int32_t model, uint32_t block_threads,
                        uint32_t blocks_per_cu);
/* Set / read the launch shape of one kernel (GS_KERNEL_*) of `model`; the
 * getter returns the shape a launch would use, defaults filled in. */
gs_status gs_set_launch_shape(gs_ctx *ctx, int32_t model, int32_t kernel, const gs_launch_shape *shape);
gs_status gs_get_launch_shape(const gs_ctx *ctx, int32_t model, int32_t kernel, gs_launch_shape *shape);

/* Order in which gs_run_plain hands particles to the lanes (results do not
 * depend on it: particles are keyed by id and sums are exact).
 * GS_ORDER_ENERGY (default): ascending injection energy for the backward
 * models, whose lifetimes fall steeply with energy, so the longest
 * trajectories start first and the run does not end with most lanes idle;
 * 1D F-p and fixed-energy runs keep id order.  GS_ORDER_ID: id order.
 * Replaces the reference's fixed id-to-thread mapping
 * (CUDAKernel/src/OneDimensionBpModel.cu:176-184: launch k integrates ids
 * [k 2^20, (k+1) 2^20) and waits for its slowest particle). */
#define GS_ORDER_ID 0
#define GS_ORDER_ENERGY 1
gs_status gs_set_order(gs_ctx *ctx, int32_t mode);
/* Autotuner: pick the launch shape of one kernel of the pilots' model by
 * timing the pilot work (GS_KERNEL_PLAIN: one simulation through
 * gs_run_plain; GS_KERNEL_BATCH: n_pilots simulations of one model through
 * gs_run_batch), best of `repeats` runs per shape, rated by particle-steps
 * per second of trajectory-kernel time.  Search: blocks per CU from 1 to the
 * occupancy at 256 threads; then 128 and 512 threads at about the same waves
 * per CU and at their occupancy; then 16, 32, 128 and 256 ids per grab.  The
 * best shape is applied (gs_set_launch_shape) and kept for gs_autotune_save.
 * best_shape / best_steps_per_s may be NULL.  With the environment variable
 * GS_AUTOTUNE_LOG set (not "0"), each measured shape and its rate is printed
 * on stderr as it completes.  Replaces the fixed launch
 * geometry chosen by compute capability in setThreadBlockSize
 * (reference CUDAKernel/src/OneDimensionBpGpuModel.cpp:45-92). */
gs_status gs_autotune_shape(gs_ctx *ctx, const gs_params *pilots, uint32_t n_pilots, int32_t kernel,
                            uint32_t repeats, gs_launch_shape *best_shape, double *best_steps_per_s);
/* gs_autotune_shape of the plain kernel with one pilot; returns its blocks per CU. */
gs_status gs_autotune(gs_ctx *ctx, const gs_params *pilot, uint32_t repeats,
                      uint32_t *best_bpcu, double *best_steps_per_s);
/* The autotuner's cache, a text file with one line per device architecture,
 * CU count, kernel and model ("<arch> <CUs> <plain|batch> <model>
 * <block_threads> <blocks_per_cu> <chunk> <steps/s>").  save writes the
 * shapes tuned on ctx, keeping the file's other lines; load applies the lines
 * that match ctx's device (*n_applied may be NULL). */
gs_status gs_autotune_save(const gs_ctx *ctx, const char *path);
gs_status gs_autotune_load(gs_ctx *ctx, const char *path, uint32_t *n_applied);
/* Statistics of the last run on ctx. */
gs_status gs_last_run_stats(const gs_ctx *ctx, gs_run_stats *st);

/* The spectrum of a run made of n parts (id-range shards of one simulation,
 * e.g. one per GPU or per call, first_particle/n_particles apart): counts,
 * statistics and exact sums added limb by limb, sum_w/sum_w2 converted with
 * correct rounding.  Equal to the spectrum of the whole run bit for bit,
 * whatever the order of the parts.  Host only: needs no context or device.
 * Parts must agree in n_bins and n_weights (GS_ERR_INVALID_ARG otherwise);
 * a carry out of the top limb counts in n_acc_overflow and returns
 * GS_ERR_OVERFLOW.  out may alias parts[0].  Replaces the reference's
 * single-process accumulation of log.dat lines by the Results pass
 * (Algorithm/src/OneDimensionBpResults.cpp:52-87), for runs split across
 * devices. */
gs_status gs_spectrum_merge(const gs_spectrum *parts, uint32_t n, gs_spectrum *out);

/* ------------------------------------------------------------------------
 * Diagnostics.  No counterpart in the reference: they expose the engine's
 * FP64 math layer and normal stream so that tests can pin them (to the ulp
 * against high-precision references, and against the oracle's stream).
 * ------------------------------------------------------------------------ */
typedef enum gs_debug_fn {
    GS_DBG_RCP = 0,       /* 1/x                                       */
    GS_DBG_RSQRT = 1,     /* 1/sqrt(x)                                 */
    GS_DBG_SQRT = 2,      /* sqrt(x), x >= 0                           */
    GS_DBG_LOG = 3,       /* ln x                                      */
    GS_DBG_EXP = 4,       /* exp x                                     */
    GS_DBG_SIN_TURN = 5,  /* sin(2 pi k 2^-32), k = (uint32) x         */
    GS_DBG_COS_TURN = 6,  /* cos(2 pi k 2^-32)                         */
    GS_DBG_LOG_U32 = 7,   /* ln((k + 1) 2^-32), k = (uint32) x         */
    GS_DBG_RAW_RCP = 8,   /* hardware v_rcp_f64 seed alone             */
    GS_DBG_RAW_RSQ = 9,   /* hardware v_rsq_f64 seed alone             */
    GS_DBG_RAW_SQRT = 10, /* hardware v_sqrt_f64 alone                 */
    GS_DBG_SIN = 11,      /* sin x, |x| < 2^20                         */
    GS_DBG_COS = 12,      /* cos x, |x| < 2^20                         */
    GS_DBG_ATAN = 13,     /* atan x                                    */
    GS_DBG_ATAN_FRAC = 14, /* atan(x[i] / x[n + i]), 0 <= x[i] <= x[n + i]:
                              x holds 2n values                          */
    GS_DBG_SHEET_F = 15,  /* 2D neutral-sheet drift factor f(sy = x[i],
                             k = x[n + i]) = atan(arg) / alphaH; 2n values */
    GS_DBG_SHEET_G = 16,  /* the same pair's tan(alphaH) / (alphaH (1 + arg^2)) */
    GS_DBG_RSQRT4 = 17,   /* x^(-1/4), x > 0 normal                    */
    GS_DBG_NORMALS = 100  /* internal: normal stream (gs_debug_normals) */
} gs_debug_fn;

/* y[i] = fn(x[i]) evaluated by the device code of the math layer. */
gs_status gs_debug_math(gs_ctx *ctx, int32_t fn, const double *x, double *y, uint64_t n);
/* The first n normals of particle `id` under `seed`, as the kernels draw them. */
gs_status gs_debug_normals(gs_ctx *ctx, uint64_t seed, uint64_t id, uint64_t n, double *out);
/* Host evaluation of the polynomial parts of the math layer (GS_DBG_LOG,
 * _EXP, _SIN_TURN, _COS_TURN, _LOG_U32); no device needed. */
gs_status gs_host_math(int32_t fn, const double *x, double *y, uint64_t n);

#ifdef __cplusplus
}
#endif

#endif /* GS_API_H */
