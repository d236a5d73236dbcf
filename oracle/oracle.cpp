// oracle.cpp - CPU FP64 restatement of the reference trajectory engine.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg as the checker / CPU baseline.  The product
// never links or calls it.
//
// What it restates (reference paths, relative to the reference root):
//   constants ............ CpuImplementations/include/Constants.hpp:63-127
//   injection grid ....... CpuImplementations/include/Constants.hpp:234-238
//   1D B-p step .......... CpuImplementations/src/OneDimensionBpCpuModel.cpp:67-141
//   1D F-p step .......... CpuImplementations/src/OneDimensionFpCpuModel.cpp:66-152
//                          (its outer loop `targetIterations < targetIterations`
//                          at :30 never runs; the particle loop is restated as
//                          the GPU path runs it)
//   2D SolarProp-like .... CpuImplementations/src/SolarPropLikeCpuModel.cpp:70-256
//   2D Geliosphere ....... CpuImplementations/src/GeliosphereCpuModel.cpp:76-407
//   weights / LIS ........ Algorithm/src/OneDimensionBpResults.cpp:51-60,
//                          OneDimensionFpResults.cpp:55-64,
//                          TwoDimensionBpResults.cpp:45-78
//
// Deliberate differences from the reference, shared with the GPU engine:
//   * noise: Philox4x32-10 keyed by the seed, counter (id, block, purpose),
//     Box-Muller in double, instead of std::mt19937 +
//     std::normal_distribution<float> (OneDimensionBpCpuModel.cpp:60-62);
//     the flag ORACLE_FLAG_FAITHFUL restores the latter, with the
//     reference's per-iteration thread schedule, for timing the CPU path;
//   * beta in full double precision (the reference takes sqrtf,
//     OneDimensionBpCpuModel.cpp:79);
//   * 2D Geliosphere draws 3 normals per step: the reference's B13 and B23
//     terms multiply identically-zero coefficients (GeliosphereCpuModel.cpp:
//     188-199, Krph = Ktph = 0) and are dropped;
//   * a per-particle step cap ends runaway trajectories (counted);
//   * the histogram is exact (integer fixed point), so the result does not
//     depend on thread count or order.
// Thresholds keep the reference's literal types (0.3f, 0.01f, ...).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- constants
// CpuImplementations/include/Constants.hpp:63-127
const double m0 = 1.67261e-27;
const double q = 1.60219e-19;
const double c = 2.99793e8;
const double Pi = 3.1415926535897932384626433832795;
const double T0 = m0 * c * c / (q * 1e9);
const double T0w = m0 * c * c;
const double omega = 2.866e-6;
const double A = 3.4;
const double konvF = 9.0e-5 / 2.0;
const double rh = 695510.0 / 150000000.0;
const double SPbins[30] = {0.01, 0.015, 0.0225, 0.03375, 0.050625,
                           0.0759375, 0.113906, 0.170859, 0.256289, 0.384434, 0.57665,
                           0.864976, 1.29746, 1.9462, 2.91929, 4.37894, 6.56841, 9.85261,
                           14.7789, 22.1684, 33.2526, 49.8789, 74.8183, 112.227, 168.341,
                           252.512, 378.768, 568.151, 852.227, 1278.34};
// Literal thresholds as typed in the reference 1D B-p model (float literals
// compared against doubles): OneDimensionBpCpuModel.cpp:116,130,136.
const double kBetaMinBp = (double)0.01f;
const double kReflectBp = (double)0.3f;
const uint64_t kDefaultMaxSteps = 100000000ull;

// ---------------------------------------------------------------- op counting
struct OpCount {
    uint64_t flops = 0, trans = 0, divs = 0, sqrts = 0;
};
thread_local OpCount *g_oc = nullptr;

// A double that counts the arithmetic it takes part in.  Operations whose
// operands are all parameters (plain doubles) are loop invariants and are
// not counted.
struct Cnt {
    double v;
    Cnt() : v(0.0) {}
    Cnt(double x) : v(x) {}
};
inline void tick(int f = 1) { if (g_oc) g_oc->flops += f; }
inline Cnt operator+(Cnt a, Cnt b) { tick(); return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { tick(); return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { tick(); return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { tick(); if (g_oc) g_oc->divs++; return Cnt(a.v / b.v); }
inline Cnt operator+(Cnt a, double b) { return a + Cnt(b); }
inline Cnt operator-(Cnt a, double b) { return a - Cnt(b); }
inline Cnt operator*(Cnt a, double b) { return a * Cnt(b); }
inline Cnt operator/(Cnt a, double b) { return a / Cnt(b); }
inline Cnt operator+(double a, Cnt b) { return Cnt(a) + b; }
inline Cnt operator-(double a, Cnt b) { return Cnt(a) - b; }
inline Cnt operator*(double a, Cnt b) { return Cnt(a) * b; }
inline Cnt operator/(double a, Cnt b) { return Cnt(a) / b; }
inline Cnt &operator+=(Cnt &a, Cnt b) { a = a + b; return a; }
inline Cnt &operator-=(Cnt &a, Cnt b) { a = a - b; return a; }
inline bool operator<(Cnt a, double b) { return a.v < b; }
inline bool operator>(Cnt a, double b) { return a.v > b; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline Cnt sqrt(Cnt a) { tick(); if (g_oc) g_oc->sqrts++; return Cnt(std::sqrt(a.v)); }
inline Cnt trans1(double v) { tick(); if (g_oc) g_oc->trans++; return Cnt(v); }
inline Cnt sin(Cnt a) { return trans1(std::sin(a.v)); }
inline Cnt cos(Cnt a) { return trans1(std::cos(a.v)); }
inline Cnt tan(Cnt a) { return trans1(std::tan(a.v)); }
inline Cnt atan(Cnt a) { return trans1(std::atan(a.v)); }
inline Cnt acos(Cnt a) { return trans1(std::acos(a.v)); }
inline Cnt pow(Cnt a, double b) { return trans1(std::pow(a.v, b)); }
inline Cnt fabs(Cnt a) { return Cnt(std::fabs(a.v)); }
inline double val(Cnt a) { return a.v; }
inline double val(double a) { return a; }

// ---------------------------------------------------------------- Philox
inline void philox10(const uint32_t in[4], const uint32_t kin[2], uint32_t out[4]) {
    uint32_t x0 = in[0], x1 = in[1], x2 = in[2], x3 = in[3];
    uint32_t k0 = kin[0], k1 = kin[1];
    for (int round = 0; round < 10; ++round) {
        if (round) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        const uint32_t y1 = (uint32_t)p1;
        const uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        const uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

const double kTwoM32 = 2.3283064365386963e-10;  // 2^-32

// Timing mode with the reference's random source (ORACLE_FLAG_FAITHFUL): the
// running thread's std::mt19937 and std::normal_distribution<float>
// (OneDimensionBpCpuModel.cpp:60-62), set by oracle_run's worker.
thread_local std::mt19937 *g_faithful_gen = nullptr;
thread_local std::normal_distribution<float> *g_faithful_dist = nullptr;

// Normal stream of one particle: block k = Philox(ctr = {id lo, id hi, k, 0},
// key = {seed lo, seed hi}) gives normals 4k..4k+3 as two Box-Muller pairs
// (x0,x1) and (x2,x3): u1 = (x + 1) 2^-32 in (0,1], u2 = x' 2^-32 in [0,1),
// z = sqrt(-2 ln u1) * (cos, sin)(2 pi u2).
struct NormalStream {
    uint32_t key[2];
    uint32_t id_lo, id_hi;
    uint32_t k = 0;
    int idx = 4;
    double buf[4];
    NormalStream(uint64_t seed, uint64_t id) {
        key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32);
        id_lo = (uint32_t)id; id_hi = (uint32_t)(id >> 32);
    }
    void refill() {
        const uint32_t ctr[4] = {id_lo, id_hi, k, 0u};
        uint32_t x[4];
        philox10(ctr, key, x);
        ++k;
        for (int j = 0; j < 2; ++j) {
            const double u1 = ((double)x[2 * j] + 1.0) * kTwoM32;
            const double u2 = (double)x[2 * j + 1] * kTwoM32;
            const double rad = std::sqrt(-2.0 * std::log(u1));
            const double ang = (2.0 * Pi) * u2;
            buf[2 * j] = rad * std::cos(ang);
            buf[2 * j + 1] = rad * std::sin(ang);
        }
        idx = 0;
    }
    double next() {
        if (g_faithful_gen) return (double)(*g_faithful_dist)(*g_faithful_gen) * scale;
        if (idx == 4) refill();
        return buf[idx++] * scale;
    }
    double scale = 1.0;
};

// Test-only flag bit of gs_params.flags: scale every normal by (1 + 2^-52),
// a last-bit perturbation used to measure how far rounding differences grow
// along a model's trajectories.
const uint32_t kOracleFlagPerturb = 0x80000000u;
// Test-only flag bit: every normal is 0, so a step is its deterministic part
// alone (advection, drifts, energy change), for checks of the step's terms.
const uint32_t kOracleFlagNoNoise = 0x40000000u;

inline double noise_scale(const gs_params &P) {
    if (P.flags & kOracleFlagNoNoise) return 0.0;
    return (P.flags & kOracleFlagPerturb) ? 1.0 + 0x1p-52 : 1.0;
}

double injection_uniform(uint64_t seed, uint64_t id) {
    const uint32_t ctr[4] = {(uint32_t)id, (uint32_t)(id >> 32), 0u, 1u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    philox10(ctr, key, x);
    const uint64_t u53 = ((uint64_t)x[0] << 21) | (uint64_t)(x[1] >> 11);
    return ((double)u53 + 0.5) * 1.1102230246251565e-16;  // 2^-53
}

double injection_energy(const gs_params &P, uint64_t id) {
    switch (P.injection) {
    case GS_INJ_LOG_UNIFORM: {
        const double u = injection_uniform(P.seed, id);
        return P.T_min * std::exp(u * std::log(P.T_max / P.T_min));
    }
    case GS_INJ_GRID: {
        // Constants.hpp:234-238: from + (id % n) * ((to - from) / n)
        const double step = (P.T_max - P.T_min) / P.grid_n;
        return P.T_min + ((double)(id % P.grid_n) * step);
    }
    case GS_INJ_SOLARPROP_BINS:
        return SPbins[id % 30];
    default:
        return P.T_min;
    }
}

uint64_t step_cap(const gs_params &P) { return P.max_steps ? P.max_steps : kDefaultMaxSteps; }

// ---------------------------------------------------------------- models
struct Emitter {
    std::vector<gs_exit_record> *out;
    void operator()(double tinj, double texit, double r, double theta, double sumac,
                    uint64_t id) const {
        gs_exit_record e;
        e.t_inj = tinj; e.t_exit = texit; e.r = r; e.theta = theta; e.sumac = sumac; e.id = id;
        out->push_back(e);
    }
};

struct ParticleEnd {
    uint64_t steps = 0;
    int status = 0;  // 0 normal, 1 NaN, 2 step cap
};

// Last state of the most recent particle on this thread (r, theta, T, p or
// rigidity as the model keeps it, sumac), for the deterministic-limit tests.
thread_local double g_probe[5];
template <class R>
inline void probe(R r, R theta, R T, R p, R sumac) {
    g_probe[0] = val(r); g_probe[1] = val(theta); g_probe[2] = val(T);
    g_probe[3] = val(p); g_probe[4] = val(sumac);
}

// 1D B-p: OneDimensionBpCpuModel.cpp:67-141 (V enters with a minus sign,
// Constants.hpp:213-218).
template <class R>
ParticleEnd particle_bp(const gs_params &P, uint64_t id, const Emitter &emit) {
    OpCount *const oc_saved = g_oc;  // count the step loop only
    g_oc = nullptr;
    using std::sqrt;
    const double V = P.V * (-1.0);
    const double dt = P.dt, K0 = P.K0;
    const double Tinj = injection_energy(P, id);
    NormalStream ns(P.seed, id);
    ns.scale = noise_scale(P);
    const uint64_t cap = step_cap(P);
    ParticleEnd end;
    R Tkin = Tinj;
    R Rig = sqrt(Tkin * (Tkin + (2.0 * T0)));
    R p = Rig * 1e9 * q / c;
    R r = P.r_inj;
    R beta, dr, pp, dp, Kdiff;
    g_oc = oc_saved;
    while (r < 100.0002) {
        if (end.steps >= cap) { end.status = 2; { probe(r, R(0.0), Tkin, p, R(0.0)); return end; } }
        beta = sqrt(Tkin * (Tkin + T0 + T0)) / (Tkin + T0);
        Rig = (p * c / q) / 1e9;
        pp = p;
        dp = (2.0 * V * pp * dt / (3.0 * r));
        p -= dp;
        Kdiff = K0 * beta * Rig;
        const R arnum = ns.next();
        dr = (V + (2.0 * Kdiff / r)) * dt + (arnum * sqrt(2.0 * Kdiff * dt));
        r += dr;
        Rig = p * c / q;
        Tkin = (sqrt((T0 * T0 * q * q * 1e9 * 1e9) + (q * q * Rig * Rig)) - (T0 * q * 1e9)) / (q * 1e9);
        beta = sqrt(Tkin * (Tkin + T0 + T0)) / (Tkin + T0);
        end.steps++;
        if (beta > kBetaMinBp && Tkin < 200.0) {
            if ((r > 100.0) && ((r - dr) < 100.0)) {
                emit(Tinj, val(Tkin), val(r), 0.0, 0.0, id);
                { probe(r, R(0.0), Tkin, p, R(0.0)); return end; }
            }
        } else if (beta < kBetaMinBp) {
            { probe(r, R(0.0), Tkin, p, R(0.0)); return end; }
        }
        if (r < kReflectBp) {
            r -= dr;
            p = pp;
        }
    }
    if (std::isnan(val(r))) end.status = 1;
    { probe(r, R(0.0), Tkin, p, R(0.0)); return end; }
}

// 1D F-p: OneDimensionFpCpuModel.cpp:66-152.
template <class R>
ParticleEnd particle_fp(const gs_params &P, uint64_t id, const Emitter &emit) {
    OpCount *const oc_saved = g_oc;  // count the step loop only
    g_oc = nullptr;
    using std::sqrt;
    const double V = P.V;
    const double dt = P.dt, K0 = P.K0;
    const double Tinj = injection_energy(P, id);
    NormalStream ns(P.seed, id);
    ns.scale = noise_scale(P);
    const uint64_t cap = step_cap(P);
    ParticleEnd end;
    R Tkin = Tinj;
    R Tkinw = Tkin * 1e9 * q;
    R Rig = sqrt(Tkinw * (Tkinw + (2.0 * T0w))) / q;
    R p = Rig * q / c;
    R sumac = 0.0;
    R r = 100.0001;
    R beta, Kdiff, dr, dp, cfactor, rp, pp;
    g_oc = oc_saved;
    while (r < 100.0002) {
        if (end.steps >= cap) { end.status = 2; { probe(r, R(0.0), Tkin, p, sumac); return end; } }
        beta = sqrt(Tkin * (Tkin + T0 + T0)) / (Tkin + T0);
        Rig = sqrt(Tkin * (Tkin + (2.0 * T0)));
        Kdiff = K0 * beta * Rig;
        const R arnum = ns.next();
        dr = (V + (2.0 * Kdiff / r)) * dt;
        dr = dr + (arnum * sqrt(2.0 * Kdiff * dt));
        dp = 2.0 * V * p * dt / (3.0 * r);
        cfactor = 4.0 * V / (3.0 * r);
        sumac = sumac + (cfactor * dt);
        rp = r;
        pp = p;
        r = r + dr;
        p = p - dp;
        Rig = p * c / q;
        Tkin = sqrt((T0 * T0 * q * q * 1e9 * 1e9) + (q * q * Rig * Rig)) - (T0 * q * 1e9);
        Tkin = Tkin / (q * 1e9);
        end.steps++;
        if ((beta > 0.01) & (Tkin < 100.0)) {
            if ((r - 1.0) / ((r - dr) - 1.0) < 0.0) {
                emit(Tinj, val(Tkin), val(r), 0.0, val(sumac), id);
            }
        }
        if (beta < 0.01) { probe(r, R(0.0), Tkin, p, sumac); return end; }
        if (r < 0.1) {
            r = rp;
            p = pp;
        }
    }
    if (std::isnan(val(r))) end.status = 1;
    { probe(r, R(0.0), Tkin, p, sumac); return end; }
}

// Coefficients of the 2D SolarProp-like step: SolarPropLikeCpuModel.cpp:98-214.
template <class R>
struct SpCoef {
    R beta, Rig, r2, gamma, gamma2, onePlusGamma, onePlusGamma2, Bfactor, Kpar, Kper, Krr,
        Ktt, dKrrTmp, dKrr, alfa, Bfield, Larmor, alphaH, arg, f, DriftR, DriftTheta, fprime,
        DriftSheetR, sinT, cosT;
};

template <class R>
void sp_coef(const gs_params &P, R r, R theta, R Tkin, SpCoef<R> &k) {
    using std::sqrt; using std::sin; using std::cos; using std::tan; using std::atan;
    using std::acos; using std::pow;
    const double V = P.V, K0 = P.K0, ratio = P.ratio, alphaM = P.tilt, polarity = P.polarity;
    k.beta = sqrt(Tkin * (Tkin + T0 + T0)) / (Tkin + T0);
    k.Rig = sqrt(Tkin * (Tkin + (2.0 * T0)));
    k.r2 = r * r;
    k.gamma = (r * omega) * sin(theta) / V;
    k.gamma2 = k.gamma * k.gamma;
    k.onePlusGamma = 1.0 + k.gamma2;
    k.onePlusGamma2 = k.onePlusGamma * k.onePlusGamma;
    k.Bfactor = (5.0 / 3.4) * k.r2 / sqrt(k.onePlusGamma);
    if (k.Rig < 0.1)
        k.Kpar = K0 * k.beta * 0.1 * k.Bfactor / 3.0;
    else
        k.Kpar = K0 * k.beta * k.Rig * k.Bfactor / 3.0;
    k.Kper = ratio * k.Kpar;
    k.Krr = k.Kper + ((k.Kpar - k.Kper) / k.onePlusGamma);
    k.Ktt = k.Kper;
    k.dKrrTmp = 2.0 * r * omega * omega * sin(theta) * sin(theta) / (V * V);
    k.dKrr = ratio * K0 * k.beta * k.Rig *
             ((2.0 * r * sqrt(k.onePlusGamma)) - (k.r2 * k.dKrrTmp / (2.0 * sqrt(k.onePlusGamma)))) /
             k.onePlusGamma;
    k.dKrr = k.dKrr + ((1.0 - ratio) * K0 * k.beta * k.Rig *
                       ((2.0 * r * pow(k.onePlusGamma, 1.5)) -
                        (k.r2 * k.dKrrTmp * 3.0 * sqrt(k.onePlusGamma) / 2.0)) /
                       pow(k.onePlusGamma, 3.0));
    k.dKrr = k.dKrr * 5.0 / (3.0 * 3.4);
    k.sinT = sin(theta);
    k.cosT = cos(theta);
    k.alfa = (Tkin + T0 + T0) / (Tkin + T0);
    k.Bfield = A * sqrt(k.onePlusGamma) / (r * r);
    k.Larmor = 0.0225 * k.Rig / k.Bfield;
    k.alphaH = Pi / sin(alphaM + (2.0 * k.Larmor * Pi / (r * 180.0)));
    k.alphaH = k.alphaH - 1.0;
    k.alphaH = 1.0 / k.alphaH;
    k.alphaH = acos(k.alphaH);
    k.arg = (1. - (2. * theta / Pi)) * tan(k.alphaH);
    k.f = atan(k.arg) / k.alphaH;
    k.DriftR = polarity * konvF * (2.0 / (3.0 * A)) * k.Rig * k.beta * r * cos(theta) * k.gamma *
               k.f / (k.onePlusGamma2 * sin(theta));
    // SolarPropLikeCpuModel.cpp:202; the reference GPU kernel fixes the
    // polarity of this term to +1 (SolarPropLikeModel.cu:213).
    const double polTheta = (P.flags & GS_FLAG_SOLARPROP_GPU_DRIFT) ? 1.0 : polarity;
    k.DriftTheta = -1.0 * polTheta * konvF * (2.0 / (3.0 * A)) * k.Rig * k.beta * r * k.gamma *
                   (2.0 + (k.gamma * k.gamma)) * k.f / k.onePlusGamma2;
    k.fprime = 1.0 + (k.arg * k.arg);
    k.fprime = tan(k.alphaH) / k.fprime;
    k.fprime = -1.0 * k.fprime * 2.0 / (Pi * k.alphaH);
    k.DriftSheetR = polarity * konvF * (1.0 / (3.0 * A)) * k.Rig * k.beta * r * k.gamma *
                    k.fprime / k.onePlusGamma;
}

// 2D SolarProp-like: SolarPropLikeCpuModel.cpp:70-256.  The step's order of
// evaluation (sin/cos calls included) follows the reference; sp_coef holds
// the coefficient block so tests can evaluate it on its own.
template <class R>
ParticleEnd particle_sp(const gs_params &P, uint64_t id, const Emitter &emit) {
    OpCount *const oc_saved = g_oc;  // count the step loop only
    g_oc = nullptr;
    using std::sqrt; using std::sin; using std::cos; using std::fabs;
    const double V = P.V, dt = P.dt;
    const double Tinj = injection_energy(P, id);
    NormalStream ns(P.seed, id);
    ns.scale = noise_scale(P);
    const uint64_t cap = step_cap(P);
    ParticleEnd end;
    R Tkin = Tinj;
    R r = P.r_inj;
    R theta = P.theta_inj;
    SpCoef<R> k;
    R dr, dtheta, dTkin;
    g_oc = oc_saved;
    while (r < 100.0) {
        if (end.steps >= cap) { end.status = 2; { probe(r, theta, Tkin, R(0.0), R(0.0)); return end; } }
        sp_coef<R>(P, r, theta, Tkin, k);
        dr = ((-1.0 * V) + (2.0 * k.Krr / r) + k.dKrr) * dt;
        dr = dr + (R(ns.next()) * sqrt(2.0 * k.Krr * dt));
        dtheta = (k.Ktt * cos(theta)) / (k.r2 * sin(theta));
        dtheta = dtheta * dt / k.onePlusGamma;
        dtheta = dtheta + ((R(ns.next()) * sqrt(2.0 * k.Ktt * dt)) / r);
        dTkin = -2.0 * V * k.alfa * Tkin * dt / (3.0 * r);
        dr = dr + ((k.DriftR + k.DriftSheetR) * dt);
        dtheta = dtheta + (k.DriftTheta * dt / r);
        r = r + dr;
        theta = theta + dtheta;
        if (theta < 0.0) theta = fabs(theta);
        if (theta > 2.0 * Pi)
            theta = theta - (2.0 * Pi);
        else if (theta > Pi)
            theta = (2.0 * Pi) - theta;
        Tkin = Tkin - dTkin;
        end.steps++;
        if (r < 0.0) r = -1.0 * r;
        if (r > 100.0) {
            emit(Tinj, val(Tkin), val(r), val(theta), 0.0, id);
            { probe(r, theta, Tkin, R(0.0), R(0.0)); return end; }
        }
    }
    if (std::isnan(val(r))) end.status = 1;
    { probe(r, theta, Tkin, R(0.0), R(0.0)); return end; }
}

// Coefficients of the 2D Geliosphere step: GeliosphereCpuModel.cpp:100-366.
template <class R>
struct GeCoef {
    R beta, Rig, r2, delta, deltarh, deltarh2, gamma, gamma2, Cb, Cb2, Bfactor, Kpar, Kper, Krr,
        Ktt, Krt, B11Temp, B11, B12, B22, COmega, dKper, dKrr, CKtt, dKtt1, dKtt2, dKtt3, dKtt4,
        dKtt, sin2, sin3, dKrtr, dKrtt, alfa, Bfield, Larmor, alphaH, arg, f, DriftR, DriftTheta,
        fprime, DriftSheetR;
};

template <class R>
void ge_coef(const gs_params &P, R r, R theta, R Tkin, GeCoef<R> &k) {
    using std::sqrt; using std::sin; using std::cos; using std::tan; using std::atan;
    using std::acos; using std::pow;
    const double V = P.V, K0 = P.K0, ratio = P.ratio, delta0 = P.delta0, alphaM = P.tilt,
                 polarity = P.polarity;
    k.beta = sqrt(Tkin * (Tkin + T0 + T0)) / (Tkin + T0);
    k.Rig = sqrt(Tkin * (Tkin + (2.0 * T0)));
    k.r2 = r * r;
    if (theta < (1.7 * Pi / 180.) || theta > (178.3 * Pi / 180.0))
        k.delta = 0.003;
    else
        k.delta = delta0 / sin(theta);
    k.deltarh = k.delta / rh;
    k.deltarh2 = k.deltarh * k.deltarh;
    k.gamma = (r * omega) * sin(theta) / V;
    k.gamma2 = k.gamma * k.gamma;
    k.Cb = 1.0 + k.gamma2 + (k.r2 * k.deltarh2);
    k.Cb2 = k.Cb * k.Cb;
    k.Bfactor = (5. / 3.4) * k.r2 / sqrt(k.Cb);
    if (k.Rig < 0.1)
        k.Kpar = K0 * k.beta * 0.1 * k.Bfactor / 3.0;
    else
        k.Kpar = K0 * k.beta * k.Rig * k.Bfactor / 3.0;
    k.Kper = ratio * k.Kpar;
    k.Krr = k.Kper + ((k.Kpar - k.Kper) / k.Cb);
    k.Ktt = k.Kper + (k.r2 * k.deltarh2 * (k.Kpar - k.Kper) / k.Cb);
    k.Krt = k.deltarh * (k.Kpar - k.Kper) * r / k.Cb;
    // Eq. 16-20 with Krph = Ktph = 0 and Kphph = 1 (GeliosphereCpuModel.cpp:163-199):
    // every dropped product is an exact zero.
    k.B11Temp = (k.Krt * k.Krt) - (k.Krr * k.Ktt);
    k.B11 = 2.0 * k.B11Temp / (0.0 - k.Ktt);
    k.B11 = sqrt(k.B11);
    k.B12 = (0.0 - k.Krt) / (0.0 - k.Ktt);
    k.B12 = k.B12 * sqrt(2.0 * k.Ktt);
    k.B22 = k.Ktt;
    k.B22 = sqrt(2.0 * k.B22) / r;
    k.COmega = 2.0 * r * omega * omega * sin(theta) * sin(theta) / (V * V);
    k.COmega = k.COmega + (2.0 * r * k.deltarh2);
    k.dKper = ratio * K0 * k.beta * k.Rig *
              ((2.0 * r * sqrt(k.Cb)) - (k.r2 * k.COmega / (2.0 * sqrt(k.Cb)))) /
              (3.0 * (5.0 / 3.4) * k.Cb);
    k.dKrr = k.dKper + ((1.0 - ratio) * K0 * k.beta * k.Rig *
                        ((2.0 * r * pow(k.Cb, 1.5)) - (k.r2 * k.COmega * 3.0 * sqrt(k.Cb) / 2.0)) /
                        (3.0 * (5.0 / 3.4) * pow(k.Cb, 3.0)));
    if ((theta > (1.7 * Pi / 180.)) && (theta < (178.3 * Pi / 180.0))) {
        k.CKtt = sin(theta) * cos(theta) * (omega * omega * k.r2 / (V * V));
        k.dKtt1 = (-1.0 * ratio * K0 * k.beta * k.Rig * k.r2 * k.CKtt) /
                  (3.0 * (5.0 / 3.4) * pow(k.Cb, 1.5));
        k.dKtt2 = (1.0 - ratio) * K0 * k.beta * k.Rig * k.r2 * k.r2 * k.deltarh2;
        k.dKtt4 = 3.0 * k.CKtt / pow(k.Cb, 2.5);
        k.dKtt = k.dKtt1 - (k.dKtt2 * k.dKtt4);
    } else {
        k.sin2 = sin(theta) * sin(theta);
        k.sin3 = sin(theta) * sin(theta) * sin(theta);
        k.CKtt = sin(theta) * cos(theta) * (omega * omega * k.r2 / (V * V));
        k.CKtt = k.CKtt - (k.r2 * delta0 * delta0 * cos(theta) / (rh * rh * k.sin3));
        k.dKtt1 = (-1.0 * ratio * K0 * k.beta * k.Rig * k.r2 * k.CKtt) /
                  (3.0 * (5.0 / 3.4) * pow(k.Cb, 1.5));
        k.dKtt2 = (1.0 - ratio) * K0 * k.beta * k.Rig * k.r2 * k.r2 * delta0 * delta0 / (rh * rh);
        k.dKtt3 = -2.0 * (cos(theta) / k.sin3) / (3.0 * (5.0 / 3.4) * pow(k.Cb, 1.5));
        k.dKtt4 = 3.0 * (k.CKtt / k.sin2) / (3.0 * (5.0 / 3.4) * pow(k.Cb, 2.5));
        k.dKtt = k.dKtt1 + (k.dKtt2 * (k.dKtt3 - k.dKtt4));
    }
    k.dKrtr = (1.0 - ratio) * K0 * k.beta * k.Rig * k.deltarh * k.r2 /
              (3.0 * (5.0 / 3.4) * pow(k.Cb, 2.5));
    if ((theta > (1.7 * Pi / 180.)) && (theta < (178.3 * Pi / 180.0))) {
        k.dKrtt = (1.0 - ratio) * K0 * k.beta * k.Rig * k.r2 * r /
                  ((3.0 * (5.0 / 3.4) * rh * pow(k.Cb, 2.5)));
        k.dKrtt = -1.0 * k.dKrtt * k.delta;
        k.dKrtt = k.dKrtt * 3.0 * k.gamma2 * cos(theta) / sin(theta);
    } else {
        k.dKrtt = (1.0 - ratio) * K0 * k.beta * k.Rig * k.r2 * r /
                  ((3.0 * (5.0 / 3.4) * rh * pow(k.Cb, 2.5)));
        k.dKrtt = -1.0 * k.dKrtt * delta0 * cos(theta) / (sin(theta) * sin(theta));
        k.dKrtt = k.dKrtt * (1.0 - (2.0 * k.r2 * k.deltarh2) + (4.0 * k.gamma2));
    }
    k.alfa = (Tkin + T0 + T0) / (Tkin + T0);
    k.Bfield = A * sqrt(k.Cb) / (r * r);
    k.Larmor = 0.0225 * k.Rig / k.Bfield;
    k.alphaH = Pi / sin(alphaM + (2.0 * k.Larmor * Pi / (r * 180.0)));
    k.alphaH = k.alphaH - 1.0;
    k.alphaH = 1.0 / k.alphaH;
    k.alphaH = acos(k.alphaH);
    k.arg = (1. - (2. * theta / Pi)) * tan(k.alphaH);
    k.f = atan(k.arg) / k.alphaH;
    k.DriftR = polarity * konvF * (2.0 / (3.0 * A)) * k.Rig * k.beta * r * cos(theta) * k.gamma *
               k.f / (k.Cb2 * sin(theta));
    k.DriftTheta = -1.0 * polarity * konvF * (2.0 / (3.0 * A)) * k.Rig * k.beta * r * k.gamma *
                   (2.0 + (k.gamma * k.gamma)) * k.f / k.Cb2;
    k.fprime = 1.0 + (k.arg * k.arg);
    k.fprime = tan(k.alphaH) / k.fprime;
    k.fprime = -1.0 * k.fprime * 2.0 / (Pi * k.alphaH);
    k.DriftSheetR = polarity * konvF * (1.0 / (3.0 * A)) * k.Rig * k.beta * r * k.gamma *
                    k.fprime / k.Cb;
}

// 2D Geliosphere: GeliosphereCpuModel.cpp:76-407.
template <class R>
ParticleEnd particle_ge(const gs_params &P, uint64_t id, const Emitter &emit) {
    OpCount *const oc_saved = g_oc;  // count the step loop only
    g_oc = nullptr;
    using std::sqrt; using std::sin; using std::cos; using std::fabs;
    const double V = P.V, dt = P.dt;
    const double Tinj = injection_energy(P, id);
    NormalStream ns(P.seed, id);
    ns.scale = noise_scale(P);
    const uint64_t cap = step_cap(P);
    ParticleEnd end;
    R Tkin = Tinj;
    R r = P.r_inj;
    R theta = P.theta_inj;
    GeCoef<R> k;
    R dr, dtheta, dTkin;
    g_oc = oc_saved;
    while (r < 100.0) {
        if (end.steps >= cap) { end.status = 2; { probe(r, theta, Tkin, R(0.0), R(0.0)); return end; } }
        ge_coef<R>(P, r, theta, Tkin, k);
        dr = ((-1.0 * V) + (2.0 * k.Krr / r) + k.dKrr) * dt;
        dr = dr + (k.dKrtt * dt / r) + (k.Krt * cos(theta) * dt / (r * sin(theta)));
        dr = dr + (R(ns.next()) * k.B11 * std::sqrt(dt));
        dr = dr + (R(ns.next()) * k.B12 * std::sqrt(dt));
        dtheta = (k.Ktt * cos(theta)) / (k.r2 * sin(theta));
        dtheta = (dtheta * dt) + (k.dKtt * dt / k.r2);
        dtheta = dtheta + (k.dKrtr * dt) + (2.0 * k.Krt * dt / r);
        dtheta = dtheta + (R(ns.next()) * k.B22 * std::sqrt(dt));
        dTkin = -2.0 * V * k.alfa * Tkin * dt / (3.0 * r);
        dr = dr + ((k.DriftR + k.DriftSheetR) * dt);
        dtheta = dtheta + (k.DriftTheta * dt / r);
        r = r + dr;
        theta = theta + dtheta;
        Tkin = Tkin - dTkin;
        end.steps++;
        if (theta < 0.0) theta = fabs(theta);
        if (theta > 2.0 * Pi)
            theta = theta - (2.0 * Pi);
        else if (theta > Pi)
            theta = (2.0 * Pi) - theta;
        if (r < 0.0) r = -1.0 * r;
        if (r > 100.0) {
            emit(Tinj, val(Tkin), val(r), val(theta), 0.0, id);
            { probe(r, theta, Tkin, R(0.0), R(0.0)); return end; }
        }
    }
    if (std::isnan(val(r))) end.status = 1;
    { probe(r, theta, Tkin, R(0.0), R(0.0)); return end; }
}

template <class R>
ParticleEnd run_particle(const gs_params &P, uint64_t id, const Emitter &emit) {
    switch (P.model) {
    case GS_MODEL_1D_FP: return particle_fp<R>(P, id, emit);
    case GS_MODEL_1D_BP: return particle_bp<R>(P, id, emit);
    case GS_MODEL_2D_SOLARPROP: return particle_sp<R>(P, id, emit);
    default: return particle_ge<R>(P, id, emit);
    }
}

// ---------------------------------------------------------------- weights
// Momentum in SI from kinetic energy in GeV (ResultConstants.hpp:19-43,
// OneDimensionBpResults.cpp:51-56).
double p_si(double T) {
    const double Tkinw = T * 1e9 * q;
    const double Rig = std::sqrt(Tkinw * (Tkinw + (2 * T0w))) / q;
    return Rig * q / c;
}
double pgev2(double T) { return T * (T + 2.0 * T0); }

double j_weber(double T) {  // TwoDimensionBpResults.cpp:50-52
    const double tem5 = 21.1 * std::exp(-2.8 * std::log(T));
    const double tem6 = 1.0 + (5.85 * std::exp(-1.22 * std::log(T))) +
                        (1.18 * std::exp(-2.54 * std::log(T)));
    return tem5 / tem6;
}
double j_usoskin(double T) {  // TwoDimensionBpResults.cpp:59-65
    const double Tr = 0.938;
    const double tt = T + Tr;
    const double t2 = tt + Tr;
    const double beta = std::sqrt(T * t2) / tt;
    double tem = (T + 0.67) / 1.67;
    tem = std::exp(-3.93 * std::log(tem));
    double j = 2.7e3 * std::exp(1.12 * std::log(T)) * tem;
    return j / (beta * beta);
}
double j_burger(double T) {  // TwoDimensionBpResults.cpp:72-76
    double tem = std::sqrt((T * T) + (2.0 * T * 0.938));
    double j = 1.9 * 1.0e4 * std::exp(-2.78 * std::log(tem));
    tem = std::sqrt((T * T) + (2.0 * T * 0.938));
    tem = 0.4866 * std::exp(-2.51 * std::log(tem));
    return j / (1 + tem);
}
double j_powerlaw(double T) { return std::exp(-2.78 * std::log(T + 0.94)); }
double f_yamada(double p) {  // OneDimensionBpResults.cpp:58-59
    const double w = (m0 * m0 * c * c * c * c) + (p * p * c * c);
    return std::exp(-1.85 * std::log(w)) / p;
}

double event_weight(const gs_params &P, int kind, const gs_exit_record &e) {
    if (P.model == GS_MODEL_1D_FP) {
        // OneDimensionFpResults.cpp:55-64: w = w0(p_inj) p_1AU^2 exp(sumac),
        // zero when r < 0.3.  Log-uniform injection carries the density ratio
        // to a uniform injection on [T_min, T_max].
        if (e.r < 0.3) return 0.0;
        double jac = 1.0;
        if (P.injection == GS_INJ_LOG_UNIFORM)
            jac = e.t_inj * std::log(P.T_max / P.T_min) / (P.T_max - P.T_min);
        const double es = std::exp(e.sumac);
        switch (kind) {
        case GS_W_YAMADA: {
            const double p100 = p_si(e.t_inj), p1 = p_si(e.t_exit);
            const double w0 = f_yamada(p100) / 1e45;
            return w0 * p1 * p1 * es * jac;
        }
        case GS_W_POWERLAW:
            return j_powerlaw(e.t_inj) / pgev2(e.t_inj) * pgev2(e.t_exit) * es * jac;
        default: {
            const double pi = p_si(e.t_inj), po = p_si(e.t_exit);
            const double j = kind == GS_W_WEBER ? j_weber(e.t_inj)
                           : kind == GS_W_USOSKIN ? j_usoskin(e.t_inj) : j_burger(e.t_inj);
            return j / (pi * pi) * (po * po) * es * jac;
        }
        }
    }
    // Backward models: the particle starts at the observer with T_inj and
    // leaves the heliosphere with T_exit; w = J(T_exit)/p_exit^2 * p_inj^2.
    switch (kind) {
    case GS_W_YAMADA: {
        const double p = p_si(e.t_exit), p1 = p_si(e.t_inj);
        return f_yamada(p) * p1 * p1;
    }
    case GS_W_POWERLAW:
        return j_powerlaw(e.t_exit) * pgev2(e.t_inj) / pgev2(e.t_exit);
    default: {
        const double p = p_si(e.t_exit), p1 = p_si(e.t_inj);
        // Weber is evaluated at the injection energy, as the reference does
        // (TwoDimensionBpResults.cpp:50-52).
        const double j = kind == GS_W_WEBER ? j_weber(e.t_inj)
                       : kind == GS_W_USOSKIN ? j_usoskin(e.t_exit) : j_burger(e.t_exit);
        return (j / (p * p)) * p1 * p1;
    }
    }
}

// ---------------------------------------------------------------- exact sums
int acc_add(uint32_t *L, double v) {
    if (v == 0.0) return 0;
    if (!(v > 0.0) || std::isinf(v)) return 2;
    int e;
    const double fr = std::frexp(v, &e);
    uint64_t m = (uint64_t)std::ldexp(fr, 53);
    int b = e - 53 - GS_ACC_LSB_EXP;
    int inexact = 0;
    if (b < 0) {
        if (-b >= 64) return 1;
        const uint64_t lost = m & ((1ull << -b) - 1ull);
        m >>= -b;
        inexact = lost != 0;
        b = 0;
    }
    int limb = b / 32;
    unsigned __int128 rest = (unsigned __int128)m << (b % 32);
    uint64_t carry = 0;
    for (int i = limb; i < GS_ACC_LIMBS; ++i) {
        const uint64_t s = (uint64_t)L[i] + (uint64_t)(rest & 0xffffffffu) + carry;
        rest >>= 32;
        L[i] = (uint32_t)s;
        carry = s >> 32;
        if (rest == 0 && carry == 0) return inexact;
    }
    return 2;
}

int acc_merge(uint32_t *L, const uint32_t *M) {
    uint64_t carry = 0;
    for (int i = 0; i < GS_ACC_LIMBS; ++i) {
        const uint64_t s = (uint64_t)L[i] + M[i] + carry;
        L[i] = (uint32_t)s;
        carry = s >> 32;
    }
    return carry ? 2 : 0;
}

double acc_to_double(const uint32_t *L) {
    int h = GS_ACC_LIMBS - 1;
    while (h >= 0 && L[h] == 0) --h;
    if (h < 0) return 0.0;
    unsigned __int128 top = 0;
    int used = 0;
    for (int i = h; i >= 0 && used < 4; --i, ++used) top = (top << 32) | L[i];
    bool sticky = false;
    for (int i = h - used; i >= 0; --i)
        if (L[i]) sticky = true;
    int msb = 127;
    while (!((top >> msb) & 1)) --msb;
    int shift = msb - 52;
    uint64_t mant;
    if (shift > 0) {
        const unsigned __int128 one = 1;
        const unsigned __int128 rem = top & ((one << shift) - 1);
        const unsigned __int128 half = one << (shift - 1);
        mant = (uint64_t)(top >> shift);
        bool up = rem > half || (rem == half && (sticky || (mant & 1)));
        if (up) {
            ++mant;
            if (mant >> 53) { mant >>= 1; ++shift; }
        }
    } else {
        mant = (uint64_t)(top << (-shift));
    }
    const int e2 = shift + 32 * (h - used + 1) + GS_ACC_LSB_EXP;
    return std::ldexp((double)mant, e2);
}

struct Hist {
    uint64_t count[GS_MAX_BINS] = {};
    uint32_t acc[GS_MAX_WEIGHTS][2][GS_MAX_BINS][GS_ACC_LIMBS] = {};
    uint64_t n_particles = 0, n_events = 0, n_outside = 0, total_steps = 0, n_nan = 0,
             n_step_cap = 0, n_inexact = 0, n_overflow = 0;
};

int find_bin(const gs_params &P, double x) {
    if (!(x >= P.bin_edges[0]) || !(x < P.bin_edges[P.n_bins])) return -1;
    int lo = 0, hi = (int)P.n_bins;  // edges[lo] <= x < edges[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (x < P.bin_edges[mid]) hi = mid; else lo = mid;
    }
    return lo;
}

void add_event(const gs_params &P, Hist &h, const gs_exit_record &e) {
    h.n_events++;
    const double x = P.model == GS_MODEL_1D_FP ? e.t_exit : e.t_inj;
    const int b = find_bin(P, x);
    if (b < 0) { h.n_outside++; return; }
    h.count[b]++;
    for (uint32_t k = 0; k < P.n_weights; ++k) {
        const double w = event_weight(P, P.weights[k], e);
        for (int s = 0; s < 2; ++s) {
            const int rc = acc_add(h.acc[k][s][b], s == 0 ? w : w * w);
            if (rc == 1) h.n_inexact++;
            if (rc == 2) h.n_overflow++;
        }
    }
}

void merge_hist(Hist &dst, const Hist &src) {
    for (int b = 0; b < GS_MAX_BINS; ++b) dst.count[b] += src.count[b];
    for (int k = 0; k < GS_MAX_WEIGHTS; ++k)
        for (int s = 0; s < 2; ++s)
            for (int b = 0; b < GS_MAX_BINS; ++b)
                if (acc_merge(dst.acc[k][s][b], src.acc[k][s][b])) dst.n_overflow++;
    dst.n_particles += src.n_particles; dst.n_events += src.n_events;
    dst.n_outside += src.n_outside; dst.total_steps += src.total_steps;
    dst.n_nan += src.n_nan; dst.n_step_cap += src.n_step_cap;
    dst.n_inexact += src.n_inexact; dst.n_overflow += src.n_overflow;
}

bool params_ok(const gs_params *p) {
    if (!p) return false;
    if (p->model < 0 || p->model > 3) return false;
    if (p->n_bins < 1 || p->n_bins > GS_MAX_BINS) return false;
    if (p->n_weights < 1 || p->n_weights > GS_MAX_WEIGHTS) return false;
    for (uint32_t i = 0; i < p->n_weights; ++i)
        if (p->weights[i] < 0 || p->weights[i] > 4) return false;
    if (!(p->dt > 0) || !(p->V > 0) || !(p->K0 >= 0)) return false;
    if (p->injection == GS_INJ_GRID && p->grid_n == 0) return false;
    if (p->injection == GS_INJ_LOG_UNIFORM && !(p->T_min > 0 && p->T_max > p->T_min)) return false;
    for (uint32_t i = 0; i < p->n_bins; ++i)
        if (!(p->bin_edges[i] < p->bin_edges[i + 1])) return false;
    return true;
}

}  // namespace

// ---------------------------------------------------------------- C ABI
extern "C" {

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox10(ctr, key, out);
}

double oracle_normal(uint64_t seed, uint64_t id, uint64_t n) {
    NormalStream ns(seed, id);
    ns.k = (uint32_t)(n / 4);
    ns.refill();
    return ns.buf[n % 4];
}

double oracle_injection_uniform(uint64_t seed, uint64_t id) { return injection_uniform(seed, id); }

double oracle_injection_energy(const gs_params *p, uint64_t id) { return injection_energy(*p, id); }

double oracle_event_weight(const gs_params *p, int32_t kind, const gs_exit_record *ev) {
    return event_weight(*p, kind, *ev);
}

int64_t oracle_particle(const gs_params *p, uint64_t id, gs_exit_record *recs, uint64_t cap,
                        uint64_t *steps, int *status) {
    if (!params_ok(p)) return -1;
    std::vector<gs_exit_record> ev;
    Emitter em{&ev};
    ParticleEnd end = run_particle<double>(*p, id, em);
    for (uint64_t i = 0; i < ev.size() && i < cap; ++i) recs[i] = ev[i];
    if (steps) *steps = end.steps;
    if (status) *status = end.status;
    return (int64_t)ev.size();
}

int oracle_run(const gs_params *p, gs_spectrum *out, gs_exit_record *recs, uint64_t rec_cap,
               uint64_t *n_recs, int n_threads) {
    if (!params_ok(p) || !out) return GS_ERR_INVALID_ARG;
    const gs_params P = *p;
    int nt = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    const uint64_t N = P.n_particles;
    const uint64_t kChunk = 64;
    std::atomic<uint64_t> next{0};
    std::vector<Hist *> hists(nt);
    std::vector<std::vector<gs_exit_record>> evs(nt);
    const bool keep = recs != nullptr || n_recs != nullptr;
    auto worker = [&](int t) {
        Hist *h = new Hist();
        hists[t] = h;
        std::vector<gs_exit_record> local;
        Emitter em{&local};
        for (;;) {
            const uint64_t s = next.fetch_add(kChunk);
            if (s >= N) break;
            const uint64_t e = std::min(N, s + kChunk);
            for (uint64_t i = s; i < e; ++i) {
                local.clear();
                ParticleEnd end = run_particle<double>(P, P.first_particle + i, em);
                h->n_particles++;
                h->total_steps += end.steps;
                if (end.status == 1) h->n_nan++;
                if (end.status == 2) h->n_step_cap++;
                for (const auto &ev : local) {
                    add_event(P, *h, ev);
                    if (keep) evs[t].push_back(ev);
                }
            }
        }
    };
    if (P.flags & ORACLE_FLAG_FAITHFUL) {
        // The reference's schedule (OneDimensionBpCpuModel.cpp:27-41,63-66):
        // per iteration, hardware_concurrency() threads are spawned and
        // joined, thread t taking the 101 x 250 particles from
        // (iteration * threads + t) * 25250; each new thread seeds its own
        // generator.
        const uint64_t per = 101 * 250;
        for (int t = 0; t < nt; ++t) hists[t] = new Hist();
        for (uint64_t it = 0; it * per * (uint64_t)nt < N; ++it) {
            std::vector<std::thread> th;
            for (int t = 0; t < nt; ++t)
                th.emplace_back([&, t, it]() {
                    std::seed_seq seq{(uint32_t)P.seed, (uint32_t)(P.seed >> 32), (uint32_t)it, (uint32_t)t};
                    std::mt19937 gen(seq);
                    std::normal_distribution<float> dist(0.0f, 1.0f);
                    g_faithful_gen = &gen;
                    g_faithful_dist = &dist;
                    Hist *h = hists[t];
                    std::vector<gs_exit_record> local;
                    Emitter em{&local};
                    const uint64_t s0 = (it * (uint64_t)nt + (uint64_t)t) * per;
                    for (uint64_t i = s0; i < std::min(N, s0 + per); ++i) {
                        local.clear();
                        ParticleEnd end = run_particle<double>(P, P.first_particle + i, em);
                        h->n_particles++;
                        h->total_steps += end.steps;
                        if (end.status == 1) h->n_nan++;
                        if (end.status == 2) h->n_step_cap++;
                        for (const auto &ev : local) {
                            add_event(P, *h, ev);
                            if (keep) evs[t].push_back(ev);
                        }
                    }
                    g_faithful_gen = nullptr;
                    g_faithful_dist = nullptr;
                });
            for (auto &x : th) x.join();
        }
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(worker, t);
        for (auto &x : th) x.join();
    }
    Hist *tot = new Hist();
    for (int t = 0; t < nt; ++t) { merge_hist(*tot, *hists[t]); delete hists[t]; }

    std::memset(out, 0, sizeof(*out));
    out->n_bins = P.n_bins;
    out->n_weights = P.n_weights;
    for (uint32_t b = 0; b < P.n_bins; ++b) out->count[b] = tot->count[b];
    for (uint32_t k = 0; k < P.n_weights; ++k)
        for (int s = 0; s < 2; ++s)
            for (uint32_t b = 0; b < P.n_bins; ++b) {
                std::memcpy(out->acc[k][s][b], tot->acc[k][s][b], sizeof(tot->acc[k][s][b]));
                const double v = acc_to_double(tot->acc[k][s][b]);
                if (s == 0) out->sum_w[k][b] = v; else out->sum_w2[k][b] = v;
            }
    out->n_particles = tot->n_particles;
    out->n_events = tot->n_events;
    out->n_outside = tot->n_outside;
    out->total_steps = tot->total_steps;
    out->n_nan = tot->n_nan;
    out->n_step_cap = tot->n_step_cap;
    out->n_acc_inexact = tot->n_inexact;
    out->n_acc_overflow = tot->n_overflow;
    delete tot;

    if (keep) {
        std::vector<gs_exit_record> all;
        for (auto &v : evs) all.insert(all.end(), v.begin(), v.end());
        std::sort(all.begin(), all.end(), [](const gs_exit_record &a, const gs_exit_record &b) {
            return a.id != b.id ? a.id < b.id : a.sumac < b.sumac;
        });
        if (n_recs) *n_recs = all.size();
        if (recs)
            for (uint64_t i = 0; i < all.size() && i < rec_cap; ++i) recs[i] = all[i];
    }
    return out->n_acc_overflow ? GS_ERR_OVERFLOW : GS_OK;
}

int oracle_2d_coeffs(const gs_params *p, double r, double theta, double T, double out[16]) {
    for (int i = 0; i < 16; ++i) out[i] = 0.0;
    if (p->model == GS_MODEL_2D_SOLARPROP) {
        SpCoef<double> k;
        sp_coef<double>(*p, r, theta, T, k);
        out[0] = k.Krr; out[1] = k.Ktt; out[3] = k.dKrr;
        out[7] = std::sqrt(2.0 * k.Krr); out[9] = std::sqrt(2.0 * k.Ktt) / r;
        out[10] = k.DriftR; out[11] = k.DriftTheta; out[12] = k.DriftSheetR;
        out[13] = k.Kpar; out[14] = k.Kper; out[15] = k.onePlusGamma;
        return 0;
    }
    if (p->model == GS_MODEL_2D_GELIOSPHERE) {
        GeCoef<double> k;
        ge_coef<double>(*p, r, theta, T, k);
        out[0] = k.Krr; out[1] = k.Ktt; out[2] = k.Krt; out[3] = k.dKrr; out[4] = k.dKtt;
        out[5] = k.dKrtr; out[6] = k.dKrtt; out[7] = k.B11; out[8] = k.B12; out[9] = k.B22;
        out[10] = k.DriftR; out[11] = k.DriftTheta; out[12] = k.DriftSheetR;
        out[13] = k.Kpar; out[14] = k.Kper; out[15] = k.Cb;
        return 0;
    }
    return 1;
}

int oracle_count_flops(const gs_params *p, uint64_t n_particles, double out[5]) {
    if (!params_ok(p)) return GS_ERR_INVALID_ARG;
    OpCount oc;
    uint64_t steps = 0;
    std::vector<gs_exit_record> ev;
    Emitter em{&ev};
    for (uint64_t i = 0; i < n_particles; ++i) {
        ev.clear();
        g_oc = &oc;
        ParticleEnd end = run_particle<Cnt>(*p, p->first_particle + i, em);
        g_oc = nullptr;
        steps += end.steps;
    }
    const double s = steps ? (double)steps : 1.0;
    out[0] = oc.flops / s;
    out[1] = oc.trans / s;
    out[2] = oc.divs / s;
    out[3] = oc.sqrts / s;
    out[4] = (double)steps;
    return 0;
}

int64_t oracle_particle_state(const gs_params *p, uint64_t id, double state[5],
                              uint64_t *steps, int *status) {
    if (!params_ok(p)) return -1;
    std::vector<gs_exit_record> ev;
    Emitter em{&ev};
    ParticleEnd end = run_particle<double>(*p, id, em);
    for (int i = 0; i < 5; ++i) state[i] = g_probe[i];
    if (steps) *steps = end.steps;
    if (status) *status = end.status;
    return (int64_t)ev.size();
}

double oracle_acc_to_double(const uint32_t *limbs) { return acc_to_double(limbs); }

int oracle_acc_add(uint32_t *limbs, double v) { return acc_add(limbs, v); }

}  // extern "C"
