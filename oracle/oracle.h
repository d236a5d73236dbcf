/*
 * oracle.h - CPU FP64 restatement of the reference trajectory engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * host adapters) may call or link this.  It is used by tests/ as the parity
 * checker, by __graft_entry__.smoke() as the checker of one small GPU run,
 * and by bench.py's cpu_baseline leg as the timed CPU path.
 *
 * It follows the reference's --cpu-only path (CpuImplementations/src/<Model>CpuModel.cpp)
 * step for step in double precision, with the reference's random sources
 * replaced by the same counter-based Philox4x32-10 streams the GPU engine
 * uses, so both sides integrate the same noise for the same particle id.
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H

#include <stdint.h>
#include "../include/gs_api.h"

/* gs_params.flags bit read by the oracle only: multiply every normal by
 * (1 + 2^-52) to measure the models' sensitivity to last-bit differences. */
#define ORACLE_FLAG_PERTURB 0x80000000u
/* gs_params.flags bit read by the oracle only: every normal is 0, so a step
 * is its deterministic part alone (advection, drifts, energy change). */
#define ORACLE_FLAG_NO_NOISE 0x40000000u
/* gs_params.flags bit read by the oracle only, for timing the reference's
 * CPU path as it runs: normals from a per-thread std::mt19937 through
 * std::normal_distribution<float> (OneDimensionBpCpuModel.cpp:60-62), and
 * threads spawned and joined per iteration of 101 x 250 particles each
 * (:27-41).  Results then depend on the thread count; no parity test sets it. */
#define ORACLE_FLAG_FAITHFUL 0x20000000u

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block function (Salmon et al., SC'11). */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                          uint32_t out[4]);
/* n-th standard normal of particle `id` under `seed`. */
double oracle_normal(uint64_t seed, uint64_t id, uint64_t n);
/* Injection uniform in (0,1) of particle `id`. */
double oracle_injection_uniform(uint64_t seed, uint64_t id);
/* Injection kinetic energy [GeV] of particle `id`. */
double oracle_injection_energy(const gs_params *p, uint64_t id);

/* Whole run on n_threads host threads (0 = hardware_concurrency).
 * Same contract as gs_run_plain. */
int oracle_run(const gs_params *p, gs_spectrum *out, gs_exit_record *recs,
               uint64_t rec_cap, uint64_t *n_recs, int n_threads);

/* One particle: events written to recs (up to cap); returns the number of
 * events; *steps receives its step count, *status 0 = normal end,
 * 1 = NaN, 2 = step cap. */
int64_t oracle_particle(const gs_params *p, uint64_t id, gs_exit_record *recs,
                        uint64_t cap, uint64_t *steps, int *status);

/* Same as oracle_particle, returning the particle's last state instead of
 * its events: state = {r, theta, T, p, sumac} where p is the momentum in SI
 * (1D models) and sumac the F-p loss integral; returns the event count. */
int64_t oracle_particle_state(const gs_params *p, uint64_t id, double state[5],
                              uint64_t *steps, int *status);

/* Weight of an event for weight kind `kind` (model-dependent rule). */
double oracle_event_weight(const gs_params *p, int32_t kind,
                           const gs_exit_record *ev);

/* 2D model coefficients at (r, theta, T), in the order
 *  0 Krr, 1 Ktt, 2 Krt, 3 dKrr/dr term, 4 dKtt term, 5 dKrt/dr term,
 *  6 dKrt/dtheta term, 7 B11, 8 B12, 9 B22, 10 DriftR, 11 DriftTheta,
 *  12 DriftSheetR, 13 Kpar, 14 Kper, 15 Cb
 * (terms that a model does not have are 0). Returns 0 on success. */
int oracle_2d_coeffs(const gs_params *p, double r, double theta, double T,
                     double out[16]);

/* Algorithmic FLOPs per step, counted on n_particles trajectories:
 * out[0] = FLOPs/step (+,-,*,/,sqrt, transcendentals = 1 each),
 * out[1] = transcendentals/step, out[2] = divisions/step,
 * out[3] = square roots/step, out[4] = steps counted. */
int oracle_count_flops(const gs_params *p, uint64_t n_particles, double out[5]);

/* Correctly rounded double of an exact accumulator (GS_ACC_LIMBS limbs). */
double oracle_acc_to_double(const uint32_t *limbs);
/* Add v >= 0 to an exact accumulator; returns 1 if bits were truncated,
 * 2 on overflow, 0 otherwise. */
int oracle_acc_add(uint32_t *limbs, double v);

#ifdef __cplusplus
}
#endif
#endif
