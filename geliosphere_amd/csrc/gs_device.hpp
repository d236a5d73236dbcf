// gs_device.hpp - device-side pieces of the trajectory engine (gfx950).
//
// One lane integrates one particle at a time.  Per-particle state lives in
// VGPRs; nothing is read from or written to HBM per step.  Events (exits,
// 1 AU crossings) become 64-byte records in the simulation's event region;
// the fold kernel bins them into histograms whose weighted sums are exact
// integers (fixed point, 32-bit digits in 64-bit words), so the result does
// not depend on the order in which lanes, waves or workgroups add events.
//
// Model semantics follow the reference CPU path (thresholds, reflections,
// order of normal draws); the arithmetic is re-associated for the GPU
// (shared reciprocals, CSE of sin/cos, pow(Cb, k/2) as products of sqrt(Cb),
// the neutral-sheet drift's acos/tan/atan chain in fraction form).  sqrt_nn
// where the argument cannot be negative (T > 0: energies only grow on the
// backward paths; 1 + gamma^2; pi (pi - 2 sin); diffusion coefficients);
// sqrt_pos for B11, whose determinant can round below zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gs_api.h"
#include "gs_kernels.h"
#include "gs_math.hpp"

namespace gs {

// Physical constants of the reference (CpuImplementations/include/Constants.hpp:63-127);
// kM0, kQ, kC and kT0 are in gs_kernels.h.
constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kT0w = kM0 * kC * kC;              // [J]
constexpr double kOmega = 2.866e-6;
constexpr double kA = 3.4;
constexpr double kKonvF = 9.0e-5 / 2.0;
// Float literals of the reference 1D B-p model, compared against doubles
// (OneDimensionBpCpuModel.cpp:116,130,136).
constexpr double kBetaMinBp = (double)0.01f;
constexpr double kReflectBp = (double)0.3f;

// ------------------------------------------------------------------ RNG
// a ^ b ^ c in one instruction (gfx950 v_bitop3_b32, truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox10(uint32_t &x0, uint32_t &x1, uint32_t &x2, uint32_t &x3,
                                         uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t y0 = xor3((uint32_t)(p1 >> 32), x1, k0);
        const uint32_t y2 = xor3((uint32_t)(p0 >> 32), x3, k1);
        x0 = y0;
        x1 = (uint32_t)p1;
        x2 = y2;
        x3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Box-Muller tables (gs_bm_tables.h); every workgroup copies them into LDS
// (load_bm_tables, then a barrier) and reads them from there.
alignas(16) __constant__ double kBmTrigTable[fm::kBmTrigWords] = GS_BM_TRIG_TABLE;
alignas(16) __constant__ double kBmLogTable[fm::kBmLogWords] = GS_BM_LOG_TABLE;
// trig holds the sin/cos table followed by the atan reduction table
// (fm::kAtanRed at trig + kBmTrigWords), so that a step reaches both
// through one pointer.
struct alignas(16) BmTables {   // entries are read as 16-byte pairs
    double trig[fm::kBmTrigWords + fm::kAtanRedWords];
    double log[fm::kBmLogWords];
};
__device__ __forceinline__ void load_bm_tables(BmTables *t) {
    for (int i = (int)threadIdx.x; i < fm::kBmTrigWords; i += (int)blockDim.x) t->trig[i] = kBmTrigTable[i];
    for (int i = (int)threadIdx.x; i < fm::kAtanRedWords; i += (int)blockDim.x)
        t->trig[fm::kBmTrigWords + i] = fm::kAtanRed[i];
    for (int i = (int)threadIdx.x; i < fm::kBmLogWords; i += (int)blockDim.x) t->log[i] = kBmLogTable[i];
}

// Normals of a particle: block k = Philox(ctr = {id lo, id hi, k, 0},
// key = {seed lo, seed hi}); its words (x0,x1) and (x2,x3) give two
// Box-Muller pairs, u1 = (x+1) 2^-32 in (0,1], u2 = x' 2^-32:
// z = sqrt(-2 ln u1) (cos, sin)(2 pi u2).  Normal 4k+i is z[i].
__device__ __forceinline__ void normal_block(uint64_t pid, uint32_t k, uint32_t k0, uint32_t k1,
                                             const BmTables &bm, double z[4]) {
    uint32_t x0 = (uint32_t)pid, x1 = (uint32_t)(pid >> 32), x2 = k, x3 = 0u;
    philox10(x0, x1, x2, x3, k0, k1);
    double s, c;
    const double ra = fm::sqrt_nn(fm::m2log_u32d(x0, bm.log));   // -2 ln u1 >= 0
    fm::sincos_turnd(x1, bm.trig, s, c);
    z[0] = ra * c;
    z[1] = ra * s;
    const double rb = fm::sqrt_nn(fm::m2log_u32d(x2, bm.log));
    fm::sincos_turnd(x3, bm.trig, s, c);
    z[2] = rb * c;
    z[3] = rb * s;
}

__device__ __forceinline__ double injection_uniform(uint64_t seed, uint64_t pid) {
    uint32_t x0 = (uint32_t)pid, x1 = (uint32_t)(pid >> 32), x2 = 0u, x3 = 1u;
    philox10(x0, x1, x2, x3, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t u53 = ((uint64_t)x0 << 21) | (uint64_t)(x1 >> 11);
    return ((double)u53 + 0.5) * 0x1p-53;
}

// Injection energies: reference CpuImplementations/include/Constants.hpp:117-121,234-238.
__constant__ double kSolarPropBins[30] = {
    0.01, 0.015, 0.0225, 0.03375, 0.050625, 0.0759375, 0.113906, 0.170859, 0.256289, 0.384434,
    0.57665, 0.864976, 1.29746, 1.9462, 2.91929, 4.37894, 6.56841, 9.85261, 14.7789, 22.1684,
    33.2526, 49.8789, 74.8183, 112.227, 168.341, 252.512, 378.768, 568.151, 852.227, 1278.34};

__device__ __forceinline__ double injection_energy(const DevParams &P, uint64_t pid) {
    switch (P.injection) {
    case GS_INJ_LOG_UNIFORM:
        return P.T_min * fm::exp_fast(injection_uniform(P.seed, pid) * P.log_ratio);
    case GS_INJ_GRID:
        return P.T_min + ((double)(pid % P.grid_n) * P.grid_step);
    case GS_INJ_SOLARPROP_BINS:
        return kSolarPropBins[pid % 30];
    default:
        return P.T_min;
    }
}

// ------------------------------------------------------------------ weights
// Reference formulas: Algorithm/src/OneDimensionBpResults.cpp:51-60,
// OneDimensionFpResults.cpp:55-64, TwoDimensionBpResults.cpp:45-78;
// x^a is exp(a ln x).
__device__ __forceinline__ double powx(double x, double a) { return fm::exp_fast(a * fm::log_fast(x)); }
__device__ __forceinline__ double p_si(double T) {
    const double Tkinw = T * 1e9 * kQ;
    const double Rig = fm::sqrt_pos(Tkinw * (Tkinw + (2 * kT0w))) * (1.0 / kQ);
    return Rig * kQ * (1.0 / kC);
}
__device__ __forceinline__ double pgev2(double T) { return T * (T + 2.0 * kT0); }
__device__ __forceinline__ double j_weber(double T) {
    const double lt = fm::log_fast(T);
    const double tem5 = 21.1 * fm::exp_fast(-2.8 * lt);
    const double tem6 = 1.0 + (5.85 * fm::exp_fast(-1.22 * lt)) + (1.18 * fm::exp_fast(-2.54 * lt));
    return tem5 * fm::rcp(tem6);
}
__device__ __forceinline__ double j_usoskin(double T) {
    const double Tr = 0.938;
    const double tt = T + Tr;
    const double t2 = tt + Tr;
    const double beta2 = T * t2 * fm::rcp(tt * tt);
    const double tem = powx((T + 0.67) * (1.0 / 1.67), -3.93);
    return 2.7e3 * powx(T, 1.12) * tem * fm::rcp(beta2);
}
__device__ __forceinline__ double j_burger(double T) {
    const double lt = fm::log_fast(fm::sqrt_pos((T * T) + (2.0 * T * 0.938)));
    const double j = 1.9 * 1.0e4 * fm::exp_fast(-2.78 * lt);
    const double tem = 0.4866 * fm::exp_fast(-2.51 * lt);
    return j * fm::rcp(1 + tem);
}
__device__ __forceinline__ double j_powerlaw(double T) { return powx(T + 0.94, -2.78); }
__device__ __forceinline__ double f_yamada(double p) {
    const double w = (kM0 * kM0 * kC * kC * kC * kC) + (p * p * kC * kC);
    return powx(w, -1.85) * fm::rcp(p);
}

struct Event {
    double t_inj, t_exit, r, theta, sumac;
};

__device__ double event_weight(const DevParams &P, int kind, const Event &e) {
    if (P.model == GS_MODEL_1D_FP) {
        if (e.r < 0.3) return 0.0;
        const double jac =
            P.injection == GS_INJ_LOG_UNIFORM ? e.t_inj * P.log_ratio * fm::rcp(P.T_max - P.T_min) : 1.0;
        const double es = fm::exp_fast(e.sumac);
        if (kind == GS_W_YAMADA) {
            const double p100 = p_si(e.t_inj), p1 = p_si(e.t_exit);
            const double w0 = f_yamada(p100) * 1e-45;
            return w0 * p1 * p1 * es * jac;
        }
        if (kind == GS_W_POWERLAW)
            return j_powerlaw(e.t_inj) * fm::rcp(pgev2(e.t_inj)) * pgev2(e.t_exit) * es * jac;
        const double pi = p_si(e.t_inj), po = p_si(e.t_exit);
        const double j = kind == GS_W_WEBER ? j_weber(e.t_inj)
                       : kind == GS_W_USOSKIN ? j_usoskin(e.t_inj) : j_burger(e.t_inj);
        return j * fm::rcp(pi * pi) * (po * po) * es * jac;
    }
    if (kind == GS_W_YAMADA) {
        const double p = p_si(e.t_exit), p1 = p_si(e.t_inj);
        return f_yamada(p) * p1 * p1;
    }
    if (kind == GS_W_POWERLAW) return j_powerlaw(e.t_exit) * pgev2(e.t_inj) * fm::rcp(pgev2(e.t_exit));
    const double p = p_si(e.t_exit), p1 = p_si(e.t_inj);
    const double j = kind == GS_W_WEBER ? j_weber(e.t_inj)
                   : kind == GS_W_USOSKIN ? j_usoskin(e.t_exit) : j_burger(e.t_exit);
    return (j * fm::rcp(p * p)) * p1 * p1;
}

// ------------------------------------------------------------------ histogram

// Add v >= 0 exactly: v = m 2^e is cut into three 32-bit pieces that land in
// consecutive digits.  Digits are 64-bit, so up to 2^32 pieces can be added
// before any carry has to be propagated (done when the workgroup flushes).
__device__ __forceinline__ void acc_add(uint64_t *dig, uint64_t *stat, double v) {
    if (v == 0.0) return;
    if (!(v > 0.0) || isinf(v)) {
        atomicAdd((unsigned long long *)&stat[ST_OVERFLOW], 1ull);
        return;
    }
    const uint64_t bits = (uint64_t)__double_as_longlong(v);
    const int ex = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1);
    int e2;
    if (ex == 0) {
        e2 = -1074;
    } else {
        m |= (1ull << 52);
        e2 = ex - 1075;
    }
    int b = e2 - GS_ACC_LSB_EXP;
    if (b < 0) {
        if (-b >= 64) {
            atomicAdd((unsigned long long *)&stat[ST_INEXACT], 1ull);
            return;
        }
        const uint64_t lost = m & ((1ull << -b) - 1ull);
        m >>= -b;
        if (lost) atomicAdd((unsigned long long *)&stat[ST_INEXACT], 1ull);
        b = 0;
    }
    const int limb = b >> 5, sh = b & 31;
    const uint64_t v0 = m << sh;
    const uint64_t hi = sh ? (m >> (64 - sh)) : 0ull;
    const uint64_t d0 = v0 & 0xffffffffull, d1 = v0 >> 32, d2 = hi;
    if (limb + (d2 ? 2 : d1 ? 1 : 0) >= GS_ACC_LIMBS) {
        atomicAdd((unsigned long long *)&stat[ST_OVERFLOW], 1ull);
        return;
    }
    if (d0) atomicAdd((unsigned long long *)&dig[limb], (unsigned long long)d0);
    if (d1) atomicAdd((unsigned long long *)&dig[limb + 1], (unsigned long long)d1);
    if (d2) atomicAdd((unsigned long long *)&dig[limb + 2], (unsigned long long)d2);
}

__device__ __forceinline__ int find_bin(const DevParams &P, double x) {
    const int nb = (int)P.n_bins;
    if (!(x >= P.edges[0]) || !(x < P.edges[nb])) return -1;
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (x < P.edges[mid]) hi = mid; else lo = mid;
    }
    return lo;
}

// Fold one event into the workgroup histogram (LDS layout: count[nb] |
// digits[nw][2][nb][16] | stat[ST_COUNT_]).
__device__ void fold_one(const DevParams &P, uint64_t *lds, const DevEvent &e) {
    const int nb = (int)P.n_bins, nw = (int)P.n_weights;
    uint64_t *cnt = lds;
    uint64_t *dig = lds + nb;
    uint64_t *stat = dig + nw * 2 * nb * GS_ACC_LIMBS;
    const double x = P.model == GS_MODEL_1D_FP ? e.t_exit : e.t_inj;
    const int b = find_bin(P, x);
    if (b < 0) {
        atomicAdd((unsigned long long *)&stat[ST_OUTSIDE], 1ull);
        return;
    }
    atomicAdd((unsigned long long *)&cnt[b], 1ull);
    Event ev;
    ev.t_inj = e.t_inj; ev.t_exit = e.t_exit; ev.r = e.r; ev.theta = e.theta; ev.sumac = e.sumac;
    for (int k = 0; k < nw; ++k) {
        const double w = event_weight(P, P.weights[k], ev);
        acc_add(dig + ((k * 2 + 0) * nb + b) * GS_ACC_LIMBS, stat, w);
        acc_add(dig + ((k * 2 + 1) * nb + b) * GS_ACC_LIMBS, stat, w * w);
    }
}

// ------------------------------------------------------------------ models
// A step returns whether the particle goes on; events are handed to
// emit(t_inj, t_exit, r, theta, sumac) where they happen, so that nothing of
// an event stays live in registers on the common path.  `trig` is the
// workgroup's LDS copy of the sin/cos table (the 2D models' sin and cos).

// The 1D models carry momenta in units of T0: P = pc / (q T0 1e9), so that
// E^2 = P^2 + 1 takes the inline constant 1.0.  K = K0 beta P [GV] enters
// only as sqrt(2 K dt) = sqrt(2 K0 dt T0) w with w = sqrt(beta P)
// (Pm.sk2dt = sqrt(2 K0 dt T0)), and as 2 K dt / r = (sqrt(2 K dt))^2 / r.
// With beta = P / E the common case has w = P E^(-1/2) = P q and beta =
// P q^2, q = (P^2 + 1)^(-1/4) from one rsqrt4: one transcendental chain per
// step instead of a 1/sqrt(E^2) and a sqrt(2 K dt).
constexpr double kInvT0 = 1.0 / kT0;

// 1D B-p.  State: r [AU], P (units of T0), w = sqrt(beta P) with beta of
// the last computed T, which a reflection does not restore
// (OneDimensionBpCpuModel.cpp:108-140).  Pm.vdt = -V dt and Pm.cp =
// -(2/3) V dt carry the backward sign of V (Constants.hpp:213-218).
// T = T0 (E - 1) is formed only for an exit; the test T < 200 is
// E^2 < ((200 + T0) / T0)^2.
struct Bp {
    static constexpr int kNormals = 1;
    static constexpr bool kLatencyBound = false;
    static constexpr bool kVecParams = false;
    __device__ static void to_vgprs(StepParams &) {}
    static constexpr double kE2Exit = ((200.0 + kT0) * kInvT0) * ((200.0 + kT0) * kInvT0);
    double r, P, w, t_inj;
    __device__ __forceinline__ void save(double *s) const { s[0] = r; s[1] = P; s[2] = w; s[3] = t_inj; }
    __device__ __forceinline__ void load(const double *s) { r = s[0]; P = s[1]; w = s[2]; t_inj = s[3]; }
    __device__ __forceinline__ bool init(const StepParams &Pm, double T) {
        t_inj = T;
        const double t = T * kInvT0;
        P = fm::sqrt_nn(t * (t + 2.0));
        w = P * fm::rsqrt(t + 1.0);   // E = t + 1
        r = Pm.r_inj;
        return r < 100.0002;
    }
    template <class Emit>
    __device__ __forceinline__ bool step(const StepParams &Pm, const double *z, const double *trig, const Emit &emit) {
        const double inv_r = fm::rcp(r);
        const double sK = Pm.sk2dt * w;                                // sqrt(2 K dt)
        const double Pn = fma(-(Pm.cp * P), inv_r, P);               // p - 2 V p dt / (3 r)
        const double dr = sK * fma(sK, inv_r, z[0]) + Pm.vdt;       // (V + 2K/r) dt + z sqrt(2 K dt)
        const double rn = r + dr;
        const double E2 = fma(Pn, Pn, 1.0);
        const double q = fm::rsqrt4(E2);
        const double invE = q * q;
        const double bn = Pn * invE;
        const double wn = Pn * q;
        // Goes on while beta >= 0.01 (false for a NaN state, which ends it as
        // r < 100.0002 does in the reference loop).
        const bool alive = bn >= kBetaMinBp;
        if (rn > 100.0) {
            // at the outer boundary (no reflection possible): exit or end
            const bool out = bn > kBetaMinBp && E2 < kE2Exit && rn - dr < 100.0;
            if (out) emit(t_inj, kT0 * fma(E2, invE, -1.0), rn, 0.0, 0.0);
            r = rn;
            P = Pn;
            w = wn;
            return alive && !out && rn < 100.0002;
        }
        // r stays below 100.0002: rn <= 100.  A reflection restores the
        // previous r as r - dr and keeps P, while beta stays that of Pn; it
        // is rare, and a branch that a wave without one skips costs no
        // selects.
        const double P0 = P;
        r = rn;
        P = Pn;
        w = wn;
        if (rn < kReflectBp) {
            asm volatile("");   // emits nothing; keeps the branch from becoming selects
            r = rn - dr;
            P = P0;
            w = fm::sqrt_nn(bn * P0);
        }
        return alive;
    }
};

// 1D F-p.  State: r, P (units of T0), beta of the last computed T, w =
// sqrt(beta R) with R = that T's momentum (a reflection keeps P but not
// beta or R), sumac.  OneDimensionFpCpuModel.cpp:84-152.  Pm.vdt = V dt,
// Pm.cp = (2/3) V dt, Pm.csum = (4/3) V dt.  T is formed only for a
// crossing; the test T < 100 is E^2 < ((100 + T0) / T0)^2.
struct Fp {
    static constexpr int kNormals = 1;
    // A run lasts as long as its longest trajectory (a lone wave at the end),
    // so per-step latency, not issue slots, sets its time.  That wave would
    // wait for a scalar reload of the (register-starved) parameters in every
    // step; in vector registers they stay put.  The empty asm emits nothing:
    // it makes them vector values.
    static constexpr bool kLatencyBound = true;
    static constexpr bool kVecParams = true;
    __device__ static void to_vgprs(StepParams &p) {
        asm volatile("" : "+v"(p.cp), "+v"(p.csum), "+v"(p.sk2dt), "+v"(p.vdt));
    }
    static constexpr double kE2Cross = ((100.0 + kT0) * kInvT0) * ((100.0 + kT0) * kInvT0);
    double r, P, beta, w, sumac, t_inj;
    __device__ __forceinline__ void save(double *s) const {
        s[0] = r; s[1] = P; s[2] = beta; s[3] = w; s[4] = sumac; s[5] = t_inj;
    }
    __device__ __forceinline__ void load(const double *s) {
        r = s[0]; P = s[1]; beta = s[2]; w = s[3]; sumac = s[4]; t_inj = s[5];
    }
    __device__ __forceinline__ bool init(const StepParams &Pm, double T) {
        t_inj = T;
        const double t = T * kInvT0;
        P = fm::sqrt_nn(t * (t + 2.0));
        const double q2 = fm::rcp(t + 1.0);   // 1/E
        beta = P * q2;
        w = P * fm::rsqrt(t + 1.0);
        sumac = 0.0;
        r = 100.0001;
        return true;
    }
    template <class Emit>
    __device__ __forceinline__ bool step(const StepParams &Pm, const double *z, const double *trig, const Emit &emit) {
        const double inv_r = fm::rcp(r);
        const double sK = Pm.sk2dt * w;
        const double dr = sK * fma(sK, inv_r, z[0]) + Pm.vdt;
        const double Pn = fma(-(Pm.cp * P), inv_r, P);
        sumac = fma(Pm.csum, inv_r, sumac);
        const double rn = r + dr;
        const double E2 = fma(Pn, Pn, 1.0);
        const double q = fm::rsqrt4(E2);
        const double rb = rn - dr;
        if (fmin(rn, rb) < 1.0) {
            // a crossing of 1 AU needs one of rn, r - dr below 1:
            // (rn - 1) / ((rn - dr) - 1) < 0, without the division
            const double a = rn - 1.0, b = rb - 1.0;
            if (beta > 0.01 && E2 < kE2Cross && ((a < 0.0 && b >= 0.0) || (a > 0.0 && b < 0.0)))
                emit(t_inj, kT0 * fma(E2, q * q, -1.0), rn, 0.0, sumac);
        }
        const bool stop = beta < 0.01;
        w = Pn * q;
        beta = w * q;
        // a reflected step (rn < 0.1) keeps r and P; rn < 100.0002 holds
        // then as well, so the loop test reads the unreflected rn.  Selects,
        // not a branch: a run's time is its longest trajectory, where a
        // branch's latency counts and the selects' issue slots do not.
        const bool refl = rn < 0.1;
        r = refl ? r : rn;
        P = refl ? P : Pn;
        return !stop && rn < 100.0002;
    }
};


// 2D SolarProp-like: SolarPropLikeCpuModel.cpp:93-256 (Kappl 2016, Eq. 13-18,
// 26, 32-37, 42-44).
struct Sp {
    static constexpr int kNormals = 2;
    static constexpr bool kLatencyBound = false;
    static constexpr bool kVecParams = false;
    __device__ static void to_vgprs(StepParams &) {}
    double r, theta, T, t_inj;
    __device__ __forceinline__ void save(double *s) const { s[0] = r; s[1] = theta; s[2] = T; s[3] = t_inj; }
    __device__ __forceinline__ void load(const double *s) { r = s[0]; theta = s[1]; T = s[2]; t_inj = s[3]; }
    __device__ __forceinline__ bool init(const StepParams &Pm, double Tin) {
        t_inj = Tin;
        T = Tin;
        r = Pm.r_inj;
        theta = Pm.theta_inj;
        return r < 100.0;
    }
    template <class Emit>
    __device__ __forceinline__ bool step(const StepParams &Pm, const double *z, const double *trig, const Emit &emit) {
#pragma clang fp contract(fast)
        double s, c;
        fm::sincosd(theta, trig, s, c);
        const double Tp = T + 2.0 * kT0;
        const double Rig = fm::sqrt_nn(T * Tp);
        const double r2 = r * r;
        const double gamma = r * Pm.omega_v * s;
        const double g2 = gamma * gamma;
        const double opg = 1.0 + g2;
        double inv_E, inv_r, inv_opg, inv_s;
        fm::rcp4(T + kT0, r, opg, s, inv_E, inv_r, inv_opg, inv_s);
        const double beta = Rig * inv_E;
        const double sq = fm::sqrt_nn(opg);
        const double inv_sq = sq * inv_opg;
        const double inv_opg15 = inv_opg * inv_sq;
        const double inv_opg25 = inv_opg * inv_opg15;
        const double Rc = fmax(Rig, 0.1);   // P below 0.1 GV counts as 0.1
        const double kbc = Pm.ge_k0c * beta;   // K0 beta (5/3.4)/3
        // K_par = K0 beta max(P, 0.1) (5/3.4) r^2 / (3 sqrt(1 + gamma^2))
        const double Kpar = kbc * Rc * r2 * inv_sq;
        const double Kper = Pm.ratio * Kpar;
        const double Krr = fma(Kpar - Kper, inv_opg, Kper);
        // dK_rr/dr (SolarPropLikeCpuModel.cpp:150-163) with r^2 dKrrTmp = 2 r gamma^2:
        //   5/(3 3.4) K0 beta P r [ratio (2 sqrt(1+g^2) - g^2/sqrt(1+g^2)) / (1+g^2)
        //                          + (1 - ratio) sqrt(1+g^2) (2 (1+g^2) - 3 g^2) / (1+g^2)^3]
        // = C K0 beta P r [ratio (2 + g^2) / (1+g^2)^1.5 + (1 - ratio) (2 - g^2) / (1+g^2)^2.5],
        // since 5/(3 3.4) = (5/3.4)/3 = C.
        const double krc = kbc * Rig * r;
        const double dKrr = krc * fma(Pm.ratio * (2.0 + g2), inv_opg15, (1.0 - Pm.ratio) * (2.0 - g2) * inv_opg25);
        const double sKrr = fm::sqrt_nn(Pm.two_dt * Krr);
        const double sKtt = fm::sqrt_nn(Pm.two_dt * Kper);
        double dr = fma(fma(2.0 * Krr, inv_r, dKrr) - Pm.V, Pm.dt, z[0] * sKrr);
        const double cs = c * inv_s;
        double dth = fma(Kper * cs * inv_r * inv_r * inv_opg, Pm.dt, z[1] * sKtt * inv_r);
        const double dTkin = -Pm.cp * (Tp * inv_E) * T * inv_r;
        // drifts: sheet angle alphaM + 2 Larmor pi / (180 r), Larmor = (0.0225 / A) P r^2 / sq
        const double sy = fm::sind(fma(Rig * r * inv_sq, (0.0225 / kA) * (kPi / 90.0), Pm.tilt), trig);
        double f, g;
        fm::sheet_drift<fm::DevOps>(sy, 1.0 - theta * (2.0 / kPi), trig + fm::kBmTrigWords, f, g);
        const double rbg = Rig * beta * r * gamma * inv_opg;
        const double dc = rbg * ((kKonvF * (2.0 / (3.0 * kA))) * f) * inv_opg;
        // (DriftR + DriftSheetR) dt, with fprime = -(2/pi) g
        dr = fma(Pm.poldt, fma(dc, cs, (kKonvF * (1.0 / (3.0 * kA)) * (-2.0 / kPi)) * rbg * g), dr);
        // DriftTheta dt / r
        dth = fma(-Pm.polthdt * dc * (2.0 + g2), inv_r, dth);
        r = r + dr;
        theta = theta + dth;
        theta = fabs(theta);   // theta < 0 -> |theta| (also takes -0 to +0)
        if (theta > kPi) {
            asm volatile("");   // emits nothing; a pole crossing is rare: no selects on the common path
            theta = theta > 2.0 * kPi ? theta - (2.0 * kPi) : (2.0 * kPi) - theta;
        }
        T = T - dTkin;
        r = fabs(r);           // r < 0 -> -r
        if (r > 100.0) {
            emit(t_inj, T, r, theta, 0.0);
            return false;
        }
        return r < 100.0;
    }
};

// 2D Geliosphere: GeliosphereCpuModel.cpp:95-407 (model notebook Eq. 16-44).
// The B13 and B23 noise terms of the reference multiply exact zeros and are
// not drawn; dr takes normals 1 and 2, dtheta normal 3.
struct Ge {
    static constexpr int kNormals = 3;
    static constexpr bool kLatencyBound = false;
    static constexpr bool kVecParams = false;
    __device__ static void to_vgprs(StepParams &) {}
    double r, theta, T, t_inj;
    __device__ __forceinline__ void save(double *s) const { s[0] = r; s[1] = theta; s[2] = T; s[3] = t_inj; }
    __device__ __forceinline__ void load(const double *s) { r = s[0]; theta = s[1]; T = s[2]; t_inj = s[3]; }
    __device__ __forceinline__ bool init(const StepParams &Pm, double Tin) {
        t_inj = Tin;
        T = Tin;
        r = Pm.r_inj;
        theta = Pm.theta_inj;
        return r < 100.0;
    }
    template <class Emit>
    __device__ __forceinline__ bool step(const StepParams &Pm, const double *z, const double *trig, const Emit &emit) {
#pragma clang fp contract(fast)
        // The reference's factors are regrouped so that each per-simulation
        // product is formed once on the host (StepParams ge_*): with
        // C = (5/3.4)/3, kf = 1/(3 (5/3.4)) and kbrc = C K0 beta P,
        //   dK_per term  ratio kf K0 beta P r (2 + h) / Cb^1.5     (2 sqrt(Cb) - h/sqrt(Cb) = (2 + h)/sqrt(Cb))
        //   dK_rr term   (1 - ratio) kf K0 beta P r (2 - h) / Cb^2.5   (2 Cb - 3 h = 2 - h)
        //   s c r^2 omega^2 / V^2 = gamma^2 cos/sin
        // with h = Cb - 1; the noise coefficients carry sqrt(dt) and the
        // drifts polarity dt.
        double s, c;
        fm::sincosd(theta, trig, s, c);
        const double Tp = T + 2.0 * kT0;
        const double Rig = fm::sqrt_nn(T * Tp);
        const double r2 = r * r;
        double inv_E, inv_r, inv_s, inv_E_r;
        {
            // 1/(T + T0), 1/r, 1/s from one reciprocal
            const double Er = (T + kT0) * r;
            const double R = fm::rcp(Er * s);
            inv_s = Er * R;
            inv_E_r = s * R;       // 1/((T + T0) r)
            inv_E = r * inv_E_r;
            inv_r = (T + kT0) * inv_E_r;
        }
        const double beta = Rig * inv_E;
        const bool polar = (theta < Pm.th_lo) || (theta > Pm.th_hi);
        const bool band = (theta > Pm.th_lo) && (theta < Pm.th_hi);
        double deltarh = Pm.delta0_rh * inv_s;   // delta / rh
        if (polar) {
            asm volatile("");   // emits nothing; the polar caps are rare: no selects on the common path
            deltarh = 0.003 * (1.0 / kRh);
        }
        const double deltarh2 = deltarh * deltarh;
        const double gamma = r * Pm.omega_v * s;
        const double g2 = gamma * gamma;
        const double rd2 = r2 * deltarh2;
        const double h = rd2 + g2;                    // Cb - 1
        const double Cb = 1.0 + h;
        const double inv_Cb = fm::rcp(Cb);
        const double sqCb = fm::sqrt_nn(Cb);
        const double inv_sqCb = sqCb * inv_Cb;
        const double inv_Cb15 = inv_Cb * inv_sqCb;
        const double inv_Cb25 = inv_Cb * inv_Cb15;
        const double Rc = fmax(Rig, 0.1);   // P below 0.1 GV counts as 0.1
        const double kbc = Pm.ge_k0c * beta;          // C K0 beta
        const double Kpar = kbc * Rc * r2 * inv_sqCb;
        const double Kper = Pm.ratio * Kpar;
        const double dKc = (Kpar - Kper) * inv_Cb;    // (K_par - K_per) / Cb
        const double Krr = Kper + dKc;
        const double Ktt = fma(rd2, dKc, Kper);
        const double Krt = deltarh * r * dKc;
        // 1/Ktt and sqrt(2 Ktt dt) from one rsqrt
        const double rK = fm::rsqrt(Ktt);
        const double inv_Ktt = rK * rK;
        // times sqrt(dt); a rounding-negative argument counts as 0
        const double B11 = fm::sqrt_nn(fmax(Pm.two_dt * fma(Krr, Ktt, -(Krt * Krt)) * inv_Ktt, 0.0));
        const double B12 = (Pm.sq2dt * Krt) * rK;
        const double B22 = (Pm.sq2dt * Ktt) * rK * inv_r;
        const double kbrc = kbc * Rig;                // C K0 beta P
        const double krc = kbrc * r;
        const double kr2c = kbrc * r2;
        const double csi = c * inv_s;
        const double dKper = (Pm.ge_a1 * krc) * (2.0 + h) * inv_Cb15;
        const double dKrr = fma(Pm.ge_a2 * krc, (2.0 - h) * inv_Cb25, dKper);
        const double q = Pm.ge_a2 * kr2c * inv_Cb25;
        const double dKrtr = q * deltarh;
        // between the polar caps (the reference's band branches)
        double CKtt = g2 * csi;                       // s c r^2 omega^2 / V^2
        double dKtt = -(CKtt * kr2c) * fma(Pm.ge_a3 * rd2, inv_Cb25, Pm.ge_a1 * inv_Cb15);
        double dKrtt = -3.0 * (dKrtr * r) * CKtt;
        if (!band) {
            asm volatile("");   // emits nothing; keeps this rare branch from becoming selects
            // GeliosphereCpuModel.cpp:244-279,295-300 (delta = delta0 / sin)
            const double d2 = Pm.delta0_rh * Pm.delta0_rh;
            const double inv_s2 = inv_s * inv_s;
            const double inv_s3 = inv_s2 * inv_s;
            CKtt = CKtt - r2 * d2 * c * inv_s3;
            const double dKtt1 = -Pm.ge_a1 * kr2c * CKtt * inv_Cb15;
            const double dKtt2 = (Pm.ge_a3 * (1.0 / 3.0)) * kr2c * r2 * d2;
            const double dKtt3 = -2.0 * c * inv_s3 * kGeKf * inv_Cb15;
            const double dKtt4 = 3.0 * CKtt * inv_s2 * kGeKf * inv_Cb25;
            dKtt = dKtt1 + dKtt2 * (dKtt3 - dKtt4);
            dKrtt = -q * r * Pm.delta0_rh * c * inv_s2 * (1.0 - (2.0 * rd2) + (4.0 * g2));
        }
        double dr = fma(fma(fma(Krt, csi, 2.0 * Krr + dKrtt), inv_r, dKrr) - Pm.V, Pm.dt, fma(z[0], B11, z[1] * B12));
        double dth = fma(fma(fma(Ktt, csi, dKtt), inv_r, 2.0 * Krt) * inv_r + dKrtr, Pm.dt, z[2] * B22);
        const double dTkin = -Pm.cp * (Tp * inv_E) * T * inv_r;
        // drifts (Kappl 2016 Eq. 26, 32-37 with Cb in place of 1 + gamma^2); sheet
        // angle alphaM + 2 Larmor pi / (180 r), Larmor = (0.0225 / A) P r^2 / sqrt(Cb)
        const double sy = fm::sind(fma(Rig * r * inv_sqCb, (0.0225 / kA) * (kPi / 90.0), Pm.tilt), trig);
        double f, g;
        fm::sheet_drift<fm::DevOps>(sy, 1.0 - theta * (2.0 / kPi), trig + fm::kBmTrigWords, f, g);
        const double rbg = Rig * beta * r * gamma * inv_Cb;
        const double dc = rbg * ((kKonvF * (2.0 / (3.0 * kA))) * f) * inv_Cb;
        // (DriftR + DriftSheetR) dt, with fprime = -(2/pi) g
        dr = fma(Pm.poldt, fma(dc, csi, (kKonvF * (1.0 / (3.0 * kA)) * (-2.0 / kPi)) * rbg * g), dr);
        // DriftTheta dt / r
        dth = fma(-Pm.poldt * dc * (2.0 + g2), inv_r, dth);
        r = r + dr;
        theta = theta + dth;
        T = T - dTkin;
        theta = fabs(theta);   // theta < 0 -> |theta| (also takes -0 to +0)
        if (theta > kPi) {
            asm volatile("");   // emits nothing; a pole crossing is rare: no selects on the common path
            theta = theta > 2.0 * kPi ? theta - (2.0 * kPi) : (2.0 * kPi) - theta;
        }
        r = fabs(r);           // r < 0 -> -r
        if (r > 100.0) {
            emit(t_inj, T, r, theta, 0.0);
            return false;
        }
        return r < 100.0;
    }
};

}  // namespace gs
