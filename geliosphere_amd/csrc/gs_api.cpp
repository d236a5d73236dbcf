// gs_api.cpp - C ABI of the trajectory engine (host side).
//
// A gs_ctx owns one HIP stream, timing events and device buffers that are
// reused across runs (the reference allocates and frees ~80 MB of managed
// memory per run, CUDAKernel/src/OneDimensionBpGpuModel.cpp:16-42).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gs_api.h"
#include "gs_kernels.h"
#include "gs_math.hpp"

static_assert(sizeof(gs_params) == 680, "gs_params layout");
static_assert(sizeof(gs_exit_record) == 48, "gs_exit_record layout");

using gs::DevOut;
using gs::DevParams;

namespace {

const uint64_t kDefaultMaxSteps = 100000000ull;
// 1.7 and 178.3 degrees as the reference writes them (GeliosphereCpuModel.cpp:112,217).
const double kPi = 3.1415926535897932384626433832795;

}  // namespace

struct gs_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device buffers, grown on demand
    uint64_t *d_words = nullptr;   // queue | count | digits | stats for up to cap_sims sims
    size_t cap_sims = 0;
    gs::DevEvent *d_events = nullptr;
    uint64_t cap_events = 0;
    uint64_t *d_meta = nullptr;    // event offsets/caps/counters, id offsets, queues
    size_t cap_meta = 0;
    std::vector<uint64_t> h_meta;  // host staging of d_meta (outlives the async copy)
    DevParams *d_params = nullptr;
    size_t cap_params = 0;
    std::vector<DevParams> h_params;
    // batched runs: side streams for concurrent model groups, fork/join
    // events, and the graph of the last batch layout
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
    std::vector<uint64_t> graph_key;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t batch_event_budget = 1ull << 29;
    // plain runs: particles per launch and event records held (0: room for all)
    uint64_t plain_chunk = 1ull << 27;
    uint64_t plain_event_cap = 0;
    // plain runs: scheduling order of a chunk's particles (gs_set_order)
    int32_t order_mode = GS_ORDER_ENERGY;
    uint32_t *d_order = nullptr;        // permutation, cap_order entries
    uint64_t cap_order = 0;
    uint32_t *d_order_counts = nullptr;  // 2 * kOrderMaxBuckets words
    uint32_t *d_border = nullptr;        // batches: permutation over the virtual ids
    uint64_t cap_border = 0;
    uint32_t *d_bcounts = nullptr;       // batches: 2 * kOrderSimBuckets words per simulation
    uint64_t cap_bcounts = 0;
    gs::SpillRec *d_spill = nullptr;     // plain runs: two halves of cap_spill records
    uint64_t cap_spill = 0;
    // launch shape per [plain, batch][model]; zero fields take the defaults
    // (256 threads, the occupancy's blocks per CU, kChunk ids per grab)
    gs_launch_shape shape[2][4] = {};
    double tuned_rate[2][4] = {};   // particle-steps/s of the tuned shapes (0: not tuned)
    std::string arch;               // gcnArchName of the device
    // the device's fixed scratch array (gs_run_plain_noalloc), looked up once
    // per device when the context is created
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    gs_run_stats stats{};
    std::string err;
};

namespace {

gs_status fail(gs_ctx *ctx, gs_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return st;
}

#define HIP_TRY(ctx, call)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((ctx), e_ == hipErrorOutOfMemory ? GS_ERR_OUT_OF_MEMORY : GS_ERR_HIP, \
                        "%s failed: %s", #call, hipGetErrorString(e_));                     \
    } while (0)

constexpr size_t kWordsPerSim = 1 + GS_MAX_BINS + gs::kDigitsPerSim + gs::ST_COUNT_;
size_t words_per_sim() { return kWordsPerSim; }

bool finite(double x) { return std::isfinite(x); }

// Validation of a parameter block; msg receives the reason.
bool check_params(const gs_params *p, std::string &msg) {
    if (!p) { msg = "params is NULL"; return false; }
    if (p->model < GS_MODEL_1D_FP || p->model > GS_MODEL_2D_GELIOSPHERE) { msg = "unknown model"; return false; }
    if (p->injection < GS_INJ_LOG_UNIFORM || p->injection > GS_INJ_FIXED) { msg = "unknown injection"; return false; }
    if (p->n_bins < 1 || p->n_bins > GS_MAX_BINS) { msg = "n_bins out of range"; return false; }
    if (p->n_weights < 1 || p->n_weights > GS_MAX_WEIGHTS) { msg = "n_weights out of range"; return false; }
    for (uint32_t i = 0; i < p->n_weights; ++i)
        if (p->weights[i] < GS_W_POWERLAW || p->weights[i] > GS_W_BURGER) { msg = "unknown weight kind"; return false; }
    if (!finite(p->K0) || p->K0 < 0) { msg = "K0 must be finite and >= 0"; return false; }
    if (!finite(p->V) || !(p->V > 0)) { msg = "V must be finite and > 0"; return false; }
    if (!finite(p->dt) || !(p->dt > 0)) { msg = "dt must be finite and > 0"; return false; }
    if (!finite(p->r_inj) || !(p->r_inj > 0)) { msg = "r_inj must be finite and > 0"; return false; }
    if (!finite(p->theta_inj) || !finite(p->tilt) || !finite(p->polarity) || !finite(p->ratio) ||
        !finite(p->delta0)) { msg = "non-finite 2D parameter"; return false; }
    if (!finite(p->T_min) || !finite(p->T_max) || !(p->T_min > 0)) { msg = "T_min must be finite and > 0"; return false; }
    if (p->injection == GS_INJ_LOG_UNIFORM && !(p->T_max > p->T_min)) { msg = "T_max must exceed T_min"; return false; }
    if (p->injection == GS_INJ_GRID && (p->grid_n == 0 || !(p->T_max >= p->T_min))) { msg = "bad grid"; return false; }
    if (p->max_steps > 0xffffffffull) { msg = "max_steps above 2^32-1"; return false; }
    if (p->n_particles > (1ull << 62) || p->first_particle > (1ull << 62)) { msg = "particle range too large"; return false; }
    for (uint32_t i = 0; i <= p->n_bins; ++i)
        if (!finite(p->bin_edges[i])) { msg = "non-finite bin edge"; return false; }
    for (uint32_t i = 0; i < p->n_bins; ++i)
        if (!(p->bin_edges[i] < p->bin_edges[i + 1])) { msg = "bin edges must ascend"; return false; }
    return true;
}

DevParams derive(const gs_params &p) {
    DevParams d;
    std::memset(&d, 0, sizeof d);
    d.model = p.model; d.injection = p.injection; d.flags = p.flags;
    d.n_bins = p.n_bins; d.n_weights = p.n_weights; d.grid_n = p.grid_n;
    for (int i = 0; i < GS_MAX_WEIGHTS; ++i) d.weights[i] = p.weights[i];
    d.K0 = p.K0; d.V = p.V; d.dt = p.dt; d.r_inj = p.r_inj; d.theta_inj = p.theta_inj;
    d.tilt = p.tilt; d.polarity = p.polarity; d.ratio = p.ratio; d.delta0 = p.delta0;
    d.T_min = p.T_min; d.T_max = p.T_max;
    d.log_ratio = (p.T_max > 0 && p.T_min > 0) ? std::log(p.T_max / p.T_min) : 0.0;
    d.grid_step = p.grid_n ? (p.T_max - p.T_min) / p.grid_n : 0.0;
    d.sqdt = std::sqrt(p.dt);
    d.two_dt = 2.0 * p.dt;
    d.sk2dt = std::sqrt(d.two_dt * p.K0 * gs::kT0);   // the 1D models' momenta are in units of T0
    const double vs = p.model == GS_MODEL_1D_BP ? -p.V : p.V;
    d.vdt = vs * p.dt;
    d.cp = (2.0 / 3.0) * vs * p.dt;
    d.csum = (4.0 / 3.0) * p.V * p.dt;
    d.omega_v = 2.866e-6 / p.V;
    d.omega2_v2 = 2.866e-6 * 2.866e-6 / (p.V * p.V);
    d.th_lo = 1.7 * kPi / 180.;
    d.th_hi = 178.3 * kPi / 180.0;
    d.pol_theta = (p.flags & GS_FLAG_SOLARPROP_GPU_DRIFT) ? 1.0 : p.polarity;
    d.ge_k0c = p.K0 * gs::kGeC;
    d.ge_a1 = p.ratio * gs::kGeKf / gs::kGeC;
    d.ge_a2 = (1.0 - p.ratio) * gs::kGeKf / gs::kGeC;
    d.ge_a3 = 3.0 * (1.0 - p.ratio) / gs::kGeC;
    d.poldt = p.polarity * p.dt;
    d.polthdt = d.pol_theta * p.dt;
    d.sq2dt = std::sqrt(2.0 * p.dt);
    d.delta0_rh = p.delta0 / gs::kRh;
    d.n_particles = p.n_particles; d.first_particle = p.first_particle; d.seed = p.seed;
    d.max_steps = p.max_steps ? p.max_steps : kDefaultMaxSteps;
    for (uint32_t i = 0; i <= p.n_bins; ++i) d.edges[i] = p.bin_edges[i];
    return d;
}

gs_status ensure_sims(gs_ctx *ctx, size_t sims) {
    if (ctx->cap_sims >= sims) return GS_OK;
    if (ctx->d_words) { (void)hipFree(ctx->d_words); ctx->d_words = nullptr; ctx->cap_sims = 0; }
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_words, sims * words_per_sim() * sizeof(uint64_t)));
    ctx->cap_sims = sims;
    return GS_OK;
}

gs_status ensure_params(gs_ctx *ctx, size_t n) {
    if (ctx->cap_params >= n) return GS_OK;
    if (ctx->d_params) { (void)hipFree(ctx->d_params); ctx->d_params = nullptr; ctx->cap_params = 0; }
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_params, n * sizeof(DevParams)));
    ctx->cap_params = n;
    return GS_OK;
}

gs_status ensure_meta(gs_ctx *ctx, size_t words) {
    if (ctx->cap_meta >= words) return GS_OK;
    if (ctx->d_meta) { (void)hipFree(ctx->d_meta); ctx->d_meta = nullptr; ctx->cap_meta = 0; }
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_meta, words * sizeof(uint64_t)));
    ctx->cap_meta = words;
    return GS_OK;
}

gs_status ensure_events(gs_ctx *ctx, uint64_t n) {
    if (ctx->cap_events >= n) return GS_OK;
    if (ctx->d_events) { (void)hipFree(ctx->d_events); ctx->d_events = nullptr; ctx->cap_events = 0; }
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_events, n * sizeof(gs::DevEvent)));
    ctx->cap_events = n;
    return GS_OK;
}

gs_status ensure_order(gs_ctx *ctx, uint64_t n) {
    if (!ctx->d_order_counts)
        HIP_TRY(ctx, hipMalloc((void **)&ctx->d_order_counts, 2 * gs::kOrderMaxBuckets * sizeof(uint32_t)));
    if (ctx->cap_order >= n) return GS_OK;
    if (ctx->d_order) { (void)hipFree(ctx->d_order); ctx->d_order = nullptr; ctx->cap_order = 0; }
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_order, n * sizeof(uint32_t)));
    ctx->cap_order = n;
    return GS_OK;
}

gs_status ensure_spill(gs_ctx *ctx, uint64_t n) {
    if (ctx->cap_spill >= n) return GS_OK;
    if (ctx->d_spill) { (void)hipFree(ctx->d_spill); ctx->d_spill = nullptr; ctx->cap_spill = 0; }
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_spill, 2 * n * sizeof(gs::SpillRec)));
    ctx->cap_spill = n;
    return GS_OK;
}

gs_status ensure_batch_order(gs_ctx *ctx, uint64_t n, uint64_t count_words) {
    if (ctx->cap_border < n) {
        if (ctx->d_border) { (void)hipFree(ctx->d_border); ctx->d_border = nullptr; ctx->cap_border = 0; }
        HIP_TRY(ctx, hipMalloc((void **)&ctx->d_border, n * sizeof(uint32_t)));
        ctx->cap_border = n;
    }
    if (ctx->cap_bcounts < count_words) {
        if (ctx->d_bcounts) { (void)hipFree(ctx->d_bcounts); ctx->d_bcounts = nullptr; ctx->cap_bcounts = 0; }
        HIP_TRY(ctx, hipMalloc((void **)&ctx->d_bcounts, count_words * sizeof(uint32_t)));
        ctx->cap_bcounts = count_words;
    }
    return GS_OK;
}

// Energy buckets of the scheduling order (nb = 0: keep id order).  The
// backward models' lifetimes fall steeply with injection energy; 1D F-p
// particles are short-lived with no such trend, and a fixed injection
// energy has nothing to sort.  Buckets: the top bits of T's double
// representation (16 per binade and up), at most 512 over the run's T range.
gs::OrderKey order_key(const gs_params &p, const DevParams &d) {
    gs::OrderKey k{0, 0, 0};
    if (p.model == GS_MODEL_1D_FP) return k;
    double lo, hi;
    switch (p.injection) {
    case GS_INJ_LOG_UNIFORM: lo = p.T_min; hi = p.T_max; break;
    case GS_INJ_GRID: lo = p.T_min; hi = p.T_min + (double)(p.grid_n ? p.grid_n - 1 : 0) * d.grid_step; break;
    case GS_INJ_SOLARPROP_BINS: lo = 0.01; hi = 1278.34; break;
    default: return k;
    }
    if (!(lo > 0.0) || !(hi > lo) || !std::isfinite(hi)) return k;
    uint64_t blo, bhi;
    std::memcpy(&blo, &lo, sizeof lo);
    std::memcpy(&bhi, &hi, sizeof hi);
    uint32_t shift = 40;
    while ((bhi >> shift) - (blo >> shift) + 1 > 512) ++shift;
    const uint64_t nb = (bhi >> shift) - (blo >> shift) + 1;
    if (nb < 2) return k;
    k.shift = shift;
    k.kmin = blo >> shift;
    k.nb = (uint32_t)nb;
    return k;
}

// Layout of the word buffer for `sims` simulations.
struct Layout {
    uint64_t *queue, *count, *digits, *stats;
};
Layout layout(uint64_t *base, size_t sims) {
    Layout L;
    L.queue = base;
    L.count = L.queue + sims;
    L.digits = L.count + sims * GS_MAX_BINS;
    L.stats = L.digits + sims * gs::kDigitsPerSim;
    return L;
}

// Correctly rounded double of normalised 32-bit limbs.
double limbs_to_double(const uint32_t *L) {
    int h = GS_ACC_LIMBS - 1;
    while (h >= 0 && L[h] == 0) --h;
    if (h < 0) return 0.0;
    unsigned __int128 top = 0;
    int used = 0;
    for (int i = h; i >= 0 && used < 4; --i, ++used) top = (top << 32) | L[i];
    bool sticky = false;
    for (int i = h - used; i >= 0; --i) sticky |= L[i] != 0;
    int msb = 127;
    while (!((top >> msb) & 1)) --msb;
    int shift = msb - 52;
    uint64_t mant;
    if (shift > 0) {
        const unsigned __int128 one = 1;
        const unsigned __int128 rem = top & ((one << shift) - 1);
        const unsigned __int128 half = one << (shift - 1);
        mant = (uint64_t)(top >> shift);
        if (rem > half || (rem == half && (sticky || (mant & 1)))) {
            ++mant;
            if (mant >> 53) { mant >>= 1; ++shift; }
        }
    } else {
        mant = (uint64_t)(top << (-shift));
    }
    return std::ldexp((double)mant, shift + 32 * (h - used + 1) + GS_ACC_LSB_EXP);
}

// Convert one simulation's global words into a gs_spectrum.
void fill_spectrum(const gs_params &p, const uint64_t *count, const uint64_t *digits,
                   const uint64_t *stats, gs_spectrum *out) {
    std::memset(out, 0, sizeof *out);
    out->n_bins = p.n_bins;
    out->n_weights = p.n_weights;
    for (uint32_t b = 0; b < p.n_bins; ++b) out->count[b] = count[b];
    uint64_t overflow = stats[gs::ST_OVERFLOW];
    for (uint32_t k = 0; k < p.n_weights; ++k)
        for (int s = 0; s < 2; ++s)
            for (uint32_t b = 0; b < p.n_bins; ++b) {
                const uint64_t *d = digits + ((size_t)(k * 2 + s) * GS_MAX_BINS + b) * GS_ACC_LIMBS;
                uint32_t *L = out->acc[k][s][b];
                uint64_t carry = 0;
                for (int i = 0; i < GS_ACC_LIMBS; ++i) {
                    const unsigned __int128 v = (unsigned __int128)d[i] + carry;
                    L[i] = (uint32_t)v;
                    carry = (uint64_t)(v >> 32);
                }
                if (carry) ++overflow;
                const double v = limbs_to_double(L);
                if (s == 0) out->sum_w[k][b] = v; else out->sum_w2[k][b] = v;
            }
    out->n_particles = stats[gs::ST_PARTICLES];
    out->n_events = stats[gs::ST_EVENTS];
    out->n_outside = stats[gs::ST_OUTSIDE];
    out->total_steps = stats[gs::ST_STEPS];
    out->n_nan = stats[gs::ST_NAN];
    out->n_step_cap = stats[gs::ST_CAP];
    out->n_acc_inexact = stats[gs::ST_INEXACT];
    out->n_acc_overflow = overflow;
}

// The launch shape a trajectory kernel runs with: the context's shape for
// (kernel, model) with its zero fields filled in (256 threads; blocks per CU
// from the kernel's occupancy, and never above it; kChunk ids per grab).
struct Resolved {
    unsigned block, bpcu, chunk, grid;
};
Resolved resolve_shape(const gs_ctx *ctx, int model, bool batch) {
    const gs_launch_shape &sh = ctx->shape[batch ? 1 : 0][model];
    Resolved R;
    R.block = sh.block_threads ? sh.block_threads : (unsigned)gs::kBlockThreads;
    int occ = 0;
    if (gs_traj_occupancy(model, batch, R.block, &occ) != hipSuccess || occ <= 0) occ = 1;
    R.bpcu = sh.blocks_per_cu && (int)sh.blocks_per_cu <= occ ? sh.blocks_per_cu : (unsigned)occ;
    R.chunk = sh.chunk ? sh.chunk : (unsigned)gs::kChunk;
    R.grid = R.bpcu * (unsigned)ctx->n_cu;
    return R;
}


size_t fold_lds_bytes(uint32_t nb, uint32_t nw) {
    return (size_t)gs::lds_words((int)nb, (int)nw) * sizeof(uint64_t);
}

unsigned fold_blocks(uint64_t events) {
    const uint64_t b = (events + gs::kBlockThreads - 1) / gs::kBlockThreads;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 1024));
}

// Particles per trajectory launch: bounds the event buffer (B-p and 2D models
// produce at most one event per particle) to 2^27 x 64 B = 8.6 GB.
const uint64_t kMaxChunk = 1ull << 27;
const uint64_t kOrderMinParticles = 4096;   // smaller runs keep id order
// Tail compaction of plain runs (DevOut::spill): used for 1D F-p, whose
// lifetimes are heavy-tailed with no energy trend, and for runs of at most
// kCompactPerLane particles per lane; continuations stop handing on below
// kCompactMin particles, and after kCompactRounds rounds.
const uint64_t kCompactPerLane = 16;
const uint64_t kCompactMin = 1024;
const int kCompactRounds = 48;

uint64_t event_room(int model, uint64_t n) {
    if (model == GS_MODEL_1D_FP) return std::max<uint64_t>(n / 16, 1u << 16);
    return std::max<uint64_t>(n, 1);
}

gs_exit_record to_record(const gs::DevEvent &e) {
    gs_exit_record r;
    r.t_inj = e.t_inj; r.t_exit = e.t_exit; r.r = e.r; r.theta = e.theta; r.sumac = e.sumac; r.id = e.id;
    return r;
}

bool record_less(const gs_exit_record &a, const gs_exit_record &b) {
    return a.id != b.id ? a.id < b.id : a.sumac < b.sumac;
}

// Relative work per particle used to order a batch (heuristic only: it
// affects the order of work, never the results).  Diffusive time across
// the heliosphere ~ L^2 / (K0 dt), times the per-step cost of the model.
double est_cost(const gs_params &p) {
    static const double step_cost[4] = {1.0, 1.0, 4.0, 6.5};
    const double L = std::max(100.0 - std::min(p.r_inj, 100.0), 1.0);
    const double K = std::max(p.K0, 1e-12);
    return step_cost[p.model] * L * L / (K * p.dt);
}

const size_t kMaxBatchSims = 65535;   // fold grid.y limit

// Layout of the fixed scratch (64-bit words) of gs_run_plain_noalloc: the
// histogram words of one simulation, its parameters, the run counters, the
// energy-order counters and permutation, the event region.
constexpr size_t kScratchParamsAt = (kWordsPerSim + 31) / 32 * 32;   // 256-byte aligned
constexpr size_t kScratchMetaAt = kScratchParamsAt + (sizeof(DevParams) + 255) / 256 * 32;
static_assert(kScratchMetaAt + 64 <= gs::kScratchHeadWords, "scratch head too small");
constexpr size_t kScratchCountsAt = gs::kScratchHeadWords;
constexpr size_t kScratchPermAt = kScratchCountsAt + gs::kScratchCountWords;
constexpr size_t kScratchEventsAt = kScratchPermAt + gs::kScratchPermWords;
static_assert(kScratchEventsAt % 8 == 0, "event region 64-byte aligned");
static_assert(kScratchEventsAt + gs::kScratchEvents * 8 == gs::kScratchWords, "scratch layout");

// Per-simulation exit records wanted from a batch (all three NULL: none).
struct BatchRecs {
    gs_exit_record *const *recs;
    const uint64_t *caps;
    uint64_t *n_recs;
    bool want(uint32_t i) const { return recs && recs[i] && caps && caps[i] > 0; }
};

gs_status run_sub_batch(gs_ctx *ctx, const gs_params *sims, const std::vector<uint32_t> &part,
                        gs_spectrum *outs, std::vector<uint32_t> &redo, gs_run_stats &acc, const BatchRecs &R);

}  // namespace

namespace {

// One sub-batch: the simulations of `part` (indices into sims), grouped by
// model (groups of larger total work first, costlier particles first within
// a group), one persistent launch per <= kBatchMaxSims simulations, the model
// groups on concurrent streams, then one fold launch over all simulations.
// The launch sequence is captured into a hipGraph once per layout and
// replayed while the layout (counts, grids, buffers) stays the same.
gs_status run_sub_batch(gs_ctx *ctx, const gs_params *sims, const std::vector<uint32_t> &part,
                        gs_spectrum *outs, std::vector<uint32_t> &redo, gs_run_stats &acc, const BatchRecs &R) {
    const size_t m = part.size();
    // ---- order
    std::vector<uint32_t> order(part);
    double gcost[4] = {0, 0, 0, 0};
    for (uint32_t i : order) gcost[sims[i].model] += est_cost(sims[i]) * (double)sims[i].n_particles;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const int ma = sims[a].model, mb = sims[b].model;
        if (ma != mb) return gcost[ma] != gcost[mb] ? gcost[ma] > gcost[mb] : ma < mb;
        return est_cost(sims[a]) > est_cost(sims[b]);
    });
    // ---- launches: (model, first sim, count)
    struct Launch { int model; uint32_t base, count; unsigned grid, block, chunk; size_t off_pos; };
    std::vector<Launch> launches;
    for (size_t j = 0; j < m;) {
        size_t k = j;
        while (k < m && sims[order[k]].model == sims[order[j]].model && k - j < (size_t)gs::kBatchMaxSims) ++k;
        Launch L;
        L.model = sims[order[j]].model;
        L.base = (uint32_t)j;
        L.count = (uint32_t)(k - j);
        const Resolved R = resolve_shape(ctx, L.model, true);
        L.grid = R.grid;
        L.block = R.block;
        L.chunk = R.chunk;
        L.off_pos = 0;
        launches.push_back(L);
        j = k;
    }
    const size_t nl = launches.size();
    // ---- host staging: params, meta = ev_off[m] | ev_cap[m] | n_events[m] |
    //      offsets (count+1 per launch) | queues[nl] | cycles[nl]
    ctx->h_params.resize(m);
    uint32_t max_nb = 1, max_nw = 1;
    uint64_t total_room = 0;
    std::vector<uint64_t> room(m);
    for (size_t j = 0; j < m; ++j) {
        const gs_params &p = sims[order[j]];
        ctx->h_params[j] = derive(p);
        if (ctx->order_mode == GS_ORDER_ENERGY) ctx->h_params[j].order = order_key(p, ctx->h_params[j]);
        room[j] = event_room(p.model, p.n_particles);
        total_room += room[j];
        max_nb = std::max(max_nb, p.n_bins);
        max_nw = std::max(max_nw, p.n_weights);
    }
    const size_t off_base = 3 * m;
    size_t pos = off_base;
    for (Launch &L : launches) { L.off_pos = pos; pos += L.count + 1; }
    const size_t q_base = pos, c_base = pos + nl, meta_words = pos + 2 * nl;
    ctx->h_meta.assign(meta_words, 0);
    uint64_t eo = 0;
    for (size_t j = 0; j < m; ++j) {
        ctx->h_meta[j] = eo;
        ctx->h_meta[m + j] = room[j];
        eo += room[j];
    }
    std::vector<uint64_t> perm_at(nl), nv(nl);   // each launch's slice of the permutation
    uint64_t perm_total = 0;
    for (size_t l = 0; l < nl; ++l) {
        const Launch &L = launches[l];
        uint64_t v = 0;
        for (uint32_t t = 0; t < L.count; ++t) {
            ctx->h_meta[L.off_pos + t] = v;
            v += sims[order[L.base + t]].n_particles;
        }
        ctx->h_meta[L.off_pos + L.count] = v;
        perm_at[l] = perm_total;
        nv[l] = v;
        perm_total += v;
    }
    const bool ordered = ctx->order_mode == GS_ORDER_ENERGY;
    gs_status st = ordered ? ensure_batch_order(ctx, std::max<uint64_t>(perm_total, 1),
                                                2 * (uint64_t)gs::kOrderSimBuckets * m)
                           : GS_OK;
    if (st == GS_OK) st = ensure_sims(ctx, m);
    if (st == GS_OK) st = ensure_params(ctx, m);
    if (st == GS_OK) st = ensure_events(ctx, std::max<uint64_t>(total_room, 1));
    if (st == GS_OK) st = ensure_meta(ctx, std::max<size_t>(meta_words, 64));
    if (st != GS_OK) return st;
    const Layout W = layout(ctx->d_words, m);
    const size_t lds = fold_lds_bytes(max_nb, max_nw);
    uint64_t max_room = 1;
    for (uint64_t r : room) max_room = std::max(max_room, r);
    const unsigned fb = std::max(1u, std::min(fold_blocks(max_room), (unsigned)std::max<size_t>(1, 4096 / m)));
    // ---- graph (re)build when the layout differs from the cached one
    std::vector<uint64_t> key = {m, nl, (uint64_t)(uintptr_t)ctx->d_words, (uint64_t)(uintptr_t)ctx->d_events,
                                 (uint64_t)(uintptr_t)ctx->d_params, (uint64_t)(uintptr_t)ctx->d_meta, lds, fb,
                                 ordered, (uint64_t)(uintptr_t)ctx->d_border, (uint64_t)(uintptr_t)ctx->d_bcounts};
    for (const Launch &L : launches) {
        key.push_back((uint64_t)L.model);
        key.push_back(L.count);
        key.push_back(L.grid);
        key.push_back(L.block);
        key.push_back(L.chunk);
    }
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_params, ctx->h_params.data(), m * sizeof(DevParams),
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_meta, ctx->h_meta.data(), meta_words * sizeof(uint64_t),
                                hipMemcpyHostToDevice, ctx->stream));
    if (!ctx->graph_exec || key != ctx->graph_key) {
        if (ctx->graph_exec) { (void)hipGraphExecDestroy(ctx->graph_exec); ctx->graph_exec = nullptr; }
        ctx->graph_key.clear();
        hipStream_t s0 = ctx->stream;
        HIP_TRY(ctx, hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
        hipError_t e = gs_launch_zero(ctx->d_words, m * words_per_sim(), s0);
        if (e == hipSuccess) e = gs_launch_zero(ctx->d_meta + 2 * m, m, s0);          // n_events
        if (e == hipSuccess) e = gs_launch_zero(ctx->d_meta + q_base, 2 * nl, s0);    // queues, cycles
        if (e == hipSuccess) e = hipEventRecord(ctx->fork, s0);
        // model groups on up to four streams
        std::vector<int> group_of(nl);
        int n_groups = 0;
        for (size_t l = 0; l < nl; ++l) {
            if (l == 0 || launches[l].model != launches[l - 1].model) ++n_groups;
            group_of[l] = n_groups - 1;
        }
        bool used[3] = {false, false, false};
        for (size_t l = 0; l < nl && e == hipSuccess; ++l) {
            const int g = group_of[l] % 4;
            hipStream_t sg = g == 0 ? s0 : ctx->side[g - 1];
            if (g > 0 && !used[g - 1]) {
                e = hipStreamWaitEvent(sg, ctx->fork, 0);
                used[g - 1] = true;
                if (e != hipSuccess) break;
            }
            const Launch &L = launches[l];
            DevOut O;
            O.queue = ctx->d_meta + q_base + l;
            O.cycles = ctx->d_meta + c_base + l;
            O.order = nullptr;
            O.spill = nullptr; O.spill_count = nullptr; O.spill_cap = 0;
            O.resume = nullptr; O.resume_n = 0;
            O.chunk = L.chunk;
            if (ordered) {
                e = gs_launch_batch_order(ctx->d_params + L.base, L.count, ctx->d_meta + L.off_pos, nv[l],
                                          ctx->d_bcounts + 2 * (size_t)gs::kOrderSimBuckets * L.base,
                                          ctx->d_border + perm_at[l], sg);
                if (e != hipSuccess) break;
                O.order = ctx->d_border + perm_at[l];
            }
            O.g_count = W.count + (size_t)L.base * GS_MAX_BINS;
            O.g_digits = W.digits + (size_t)L.base * gs::kDigitsPerSim;
            O.g_stats = W.stats + (size_t)L.base * gs::ST_COUNT_;
            O.events = ctx->d_events;
            O.ev_off = ctx->d_meta + L.base;
            O.ev_cap = ctx->d_meta + m + L.base;
            O.n_events = ctx->d_meta + 2 * m + L.base;
            e = gs_launch_traj(L.model, true, ctx->d_params + L.base, L.count, ctx->d_meta + L.off_pos, O,
                               L.grid, L.block, sg);
        }
        for (int g = 0; g < 3 && e == hipSuccess; ++g) {
            if (!used[g]) continue;
            e = hipEventRecord(ctx->join[g], ctx->side[g]);
            if (e == hipSuccess) e = hipStreamWaitEvent(s0, ctx->join[g], 0);
        }
        if (e == hipSuccess) {
            DevOut O;
            O.queue = nullptr; O.cycles = nullptr; O.order = nullptr;
            O.spill = nullptr; O.spill_count = nullptr; O.spill_cap = 0;
            O.resume = nullptr; O.resume_n = 0;
            O.chunk = 0;
            O.g_count = W.count; O.g_digits = W.digits; O.g_stats = W.stats;
            O.events = ctx->d_events;
            O.ev_off = ctx->d_meta; O.ev_cap = ctx->d_meta + m; O.n_events = ctx->d_meta + 2 * m;
            e = gs_launch_fold(ctx->d_params, (uint32_t)m, O, lds, fb, s0);
        }
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s0, &graph);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            ctx->graph_exec = nullptr;
            return fail(ctx, GS_ERR_HIP, "batch graph: %s", hipGetErrorString(e));
        }
        ctx->graph_key = key;
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<uint64_t> host(m * words_per_sim());
    std::vector<uint64_t> hmeta(meta_words);
    HIP_TRY(ctx, hipMemcpyAsync(host.data(), ctx->d_words, host.size() * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(hmeta.data(), ctx->d_meta, meta_words * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    acc.kernel_ms += ms;
    const Layout H = layout(host.data(), m);
    for (size_t j = 0; j < m; ++j) {
        const uint32_t i = order[j];
        const uint64_t nev = hmeta[2 * m + j];
        if (nev > hmeta[m + j]) { redo.push_back(i); continue; }
        uint64_t *stats = H.stats + j * gs::ST_COUNT_;
        stats[gs::ST_EVENTS] = nev;
        fill_spectrum(sims[i], H.count + j * GS_MAX_BINS, H.digits + j * gs::kDigitsPerSim, stats, &outs[i]);
        if (R.n_recs) R.n_recs[i] = nev;
        if (R.want(i) && nev > 0) {
            // the simulation's event region, in gs_run_plain's record order
            std::vector<gs::DevEvent> h(nev);
            HIP_TRY(ctx, hipMemcpyAsync(h.data(), ctx->d_events + hmeta[j], nev * sizeof(gs::DevEvent),
                                        hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            std::vector<gs_exit_record> rs(nev);
            for (uint64_t e = 0; e < nev; ++e) rs[e] = to_record(h[e]);
            std::sort(rs.begin(), rs.end(), record_less);
            std::copy(rs.begin(), rs.begin() + (ptrdiff_t)std::min<uint64_t>(nev, R.caps[i]), R.recs[i]);
        }
        acc.active_cycles += stats[gs::ST_STEPS];
    }
    for (size_t l = 0; l < nl; ++l) {
        acc.lane_cycles += hmeta[c_base + l] * (uint64_t)gs::kWave * gs::kCycleSteps;
        acc.grid_blocks += launches[l].grid;
        acc.block_threads = launches[l].block;
    }
    return GS_OK;
}

// Device buffers of one plain run.  meta: [0] event offset (0), [1] event
// capacity, [2] event count, [3] cycles, [4..5] hand-on counts,
// [8 .. 8+ST_COUNT_) statistics saved before a chunk.
// own = true: the context's buffers, which grow on demand (event region,
// order permutation, hand-on records).  own = false: fixed buffers (events,
// order_counts and order of cap_events entries); no exit records, no
// hand-on, and chunks are cut to fit the fixed event region.
struct PlainBufs {
    uint64_t *words;
    DevParams *params;
    uint64_t *meta;
    bool own;
    gs::DevEvent *events;
    uint64_t cap_events;
    uint32_t *order_counts;
    uint32_t *order;
};

// One plain run, as consecutive chunks of the id range.  A chunk whose events
// overflow the event region runs again (trajectories are deterministic),
// from the statistics saved before it: with more room (own buffers), or as
// half as many particles (fixed buffers).
gs_status run_plain_core(gs_ctx *ctx, const gs_params &pp, gs_spectrum *out, gs_exit_record *recs,
                         uint64_t rec_cap, uint64_t *n_recs, const PlainBufs &B) {
    const gs_params *p = &pp;
    const auto t_start = std::chrono::steady_clock::now();
    const bool want_recs = recs != nullptr && rec_cap > 0;
    if (want_recs && !B.own) return fail(ctx, GS_ERR_INVALID_ARG, "exit records need the context's buffers");
    Layout L = layout(B.words, 1);
    HIP_TRY(ctx, gs_launch_zero(B.words, kWordsPerSim, ctx->stream));
    const Resolved RS = resolve_shape(ctx, p->model, false);
    const unsigned grid = RS.grid;
    std::vector<gs_exit_record> kept;   // records only (never with fixed buffers)
    uint64_t total_events = 0;
    double kernel_ms = 0.0;
    const uint64_t N = p->n_particles;
    DevOut O;
    O.queue = L.queue; O.g_count = L.count; O.g_digits = L.digits; O.g_stats = L.stats;
    O.events = nullptr;
    O.ev_off = B.meta; O.ev_cap = B.meta + 1; O.n_events = B.meta + 2;
    O.cycles = B.meta + 3;
    O.order = nullptr;
    O.spill = nullptr; O.spill_count = nullptr; O.spill_cap = 0;
    O.resume = nullptr; O.resume_n = 0;
    O.chunk = RS.chunk;
    const uint64_t lanes = (uint64_t)grid * RS.block;
    // largest chunk: 2^27 particles, or what the fixed event region holds
    // (B-p and 2D models emit at most one event per particle)
    const uint64_t chunk_max = B.own ? ctx->plain_chunk
                                     : (p->model == GS_MODEL_1D_FP ? B.cap_events * 16 : B.cap_events);
    // most event records a chunk may hold
    const uint64_t ev_limit = B.own ? (ctx->plain_event_cap ? ctx->plain_event_cap : ~0ull) : B.cap_events;
    DevParams hp;              // host copy of the chunk's parameters
    uint64_t hm[2] = {0, 0};   // event offset and capacity
    gs_status st = GS_OK;
    uint64_t c0 = 0;
    uint64_t size = std::min(chunk_max, N);
    bool grown = false;   // own buffers: the current chunk's region was grown once
    do {
        gs_params q = *p;
        q.first_particle = p->first_particle + c0;
        q.n_particles = size;
        hp = derive(q);
        if (B.own) {
            st = ensure_events(ctx, std::min(event_room(q.model, q.n_particles), ev_limit));
            if (st != GS_OK) return st;
        }
        gs::DevEvent *events = B.own ? ctx->d_events : B.events;
        const uint64_t cap = std::min(B.own ? ctx->cap_events : B.cap_events, ev_limit);
        HIP_TRY(ctx, hipMemcpyAsync(B.params, &hp, sizeof(DevParams), hipMemcpyHostToDevice, ctx->stream));
        const gs::OrderKey key = order_key(q, hp);
        O.order = nullptr;
        if (ctx->order_mode == GS_ORDER_ENERGY && key.nb > 0 && q.n_particles >= kOrderMinParticles) {
            uint32_t *counts = B.order_counts, *perm = B.order;
            if (B.own) {
                st = ensure_order(ctx, q.n_particles);
                if (st != GS_OK) return st;
                counts = ctx->d_order_counts;
                perm = ctx->d_order;
            }
            HIP_TRY(ctx, gs_launch_order(B.params, q.n_particles, key, counts, perm, ctx->stream));
            O.order = perm;
        }
        // statistics before the chunk, restored if it has to run again
        HIP_TRY(ctx, hipMemcpyAsync(B.meta + 8, L.stats, gs::ST_COUNT_ * sizeof(uint64_t),
                                    hipMemcpyDeviceToDevice, ctx->stream));
        hm[0] = 0;
        hm[1] = cap;
        HIP_TRY(ctx, hipMemcpyAsync(B.meta, hm, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, gs_launch_zero(L.queue, 1, ctx->stream));
        HIP_TRY(ctx, gs_launch_zero(B.meta + 2, 4, ctx->stream));   // events, cycles, hand-on counts
        O.events = events;
        const bool compact = B.own && (q.model == GS_MODEL_1D_FP || q.n_particles <= kCompactPerLane * lanes);
        if (compact) {
            st = ensure_spill(ctx, lanes);   // at most one hand-on per lane and launch
            if (st != GS_OK) return st;
        }
        O.resume = nullptr;
        O.resume_n = 0;
        O.spill = compact ? ctx->d_spill : nullptr;
        O.spill_count = B.meta + 4;
        O.spill_cap = compact ? ctx->cap_spill : 0;
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIP_TRY(ctx, gs_launch_traj(q.model, false, B.params, 1, nullptr, O, grid, RS.block, ctx->stream));
        // continuations: the handed-on particles, packed into full waves
        for (int round = 0, cur = 0; compact; ++round, cur ^= 1) {
            uint64_t pending = 0;
            HIP_TRY(ctx, hipMemcpyAsync(&pending, B.meta + 4 + cur, sizeof pending, hipMemcpyDeviceToHost,
                                        ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            pending = std::min(pending, O.spill_cap);   // slots beyond the cap were not handed on
            if (pending == 0) break;
            HIP_TRY(ctx, gs_launch_zero(L.queue, 1, ctx->stream));
            HIP_TRY(ctx, gs_launch_zero(B.meta + 4 + (cur ^ 1), 1, ctx->stream));
            O.resume = ctx->d_spill + (size_t)cur * ctx->cap_spill;
            O.resume_n = pending;
            O.spill = pending > kCompactMin && round + 1 < kCompactRounds
                          ? ctx->d_spill + (size_t)(cur ^ 1) * ctx->cap_spill : nullptr;
            O.spill_count = B.meta + 4 + (cur ^ 1);
            const unsigned g2 = (unsigned)std::min<uint64_t>(grid, (pending + RS.block - 1) / RS.block);
            HIP_TRY(ctx, gs_launch_traj(q.model, false, B.params, 1, nullptr, O, g2, RS.block, ctx->stream));
        }
        O.resume = nullptr;
        O.spill = nullptr;
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        uint64_t nev = 0;
        HIP_TRY(ctx, hipMemcpyAsync(&nev, B.meta + 2, sizeof nev, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        kernel_ms += ms;
        if (nev > cap) {
            // The region was too small (1D F-p crossings are unbounded): run
            // the chunk again from the statistics saved before it, with room
            // for its nev events where the limit allows, else as half as
            // many particles.
            HIP_TRY(ctx, hipMemcpyAsync(L.stats, B.meta + 8, gs::ST_COUNT_ * sizeof(uint64_t),
                                        hipMemcpyDeviceToDevice, ctx->stream));
            if (B.own && !grown && nev <= ev_limit) {
                st = ensure_events(ctx, std::min(nev + nev / 8, ev_limit));
                if (st != GS_OK) return st;
                grown = true;
                continue;
            }
            if (size == 1)
                return fail(ctx, GS_ERR_OVERFLOW, "particle %llu has more than %llu events, the event capacity",
                            (unsigned long long)q.first_particle, (unsigned long long)cap);
            size /= 2;
            grown = false;
            continue;
        }
        HIP_TRY(ctx, gs_launch_fold(B.params, 1, O, fold_lds_bytes(q.n_bins, q.n_weights), fold_blocks(nev),
                                    ctx->stream));
        total_events += nev;
        if (want_recs && nev > 0 && kept.size() < rec_cap) {
            std::vector<gs::DevEvent> h(nev);
            HIP_TRY(ctx, hipMemcpyAsync(h.data(), events, nev * sizeof(gs::DevEvent), hipMemcpyDeviceToHost,
                                        ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            std::vector<gs_exit_record> rs(nev);
            for (uint64_t i = 0; i < nev; ++i) rs[i] = to_record(h[i]);
            std::sort(rs.begin(), rs.end(), record_less);
            for (uint64_t i = 0; i < nev && kept.size() < rec_cap; ++i) kept.push_back(rs[i]);
        }
        c0 += size;
        size = std::min(chunk_max, N - c0);
        grown = false;
    } while (c0 < N);
    uint64_t host[kWordsPerSim];
    HIP_TRY(ctx, hipMemcpyAsync(host, B.words, sizeof host, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    Layout H = layout(host, 1);
    H.stats[gs::ST_EVENTS] = total_events;
    fill_spectrum(*p, H.count, H.digits, H.stats, out);
    if (want_recs) std::copy(kept.begin(), kept.end(), recs);
    if (n_recs) *n_recs = total_events;
    ctx->stats.kernel_ms = kernel_ms;
    ctx->stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    ctx->stats.lane_cycles = H.stats[gs::ST_CYCLES] * (uint64_t)gs::kWave * gs::kCycleSteps;
    ctx->stats.max_particle_steps = H.stats[gs::ST_MAX_STEPS];
    ctx->stats.active_cycles = H.stats[gs::ST_STEPS];
    ctx->stats.grid_blocks = grid;
    ctx->stats.block_threads = RS.block;
    if (out->n_acc_overflow) return fail(ctx, GS_ERR_OVERFLOW, "weight accumulator overflow");
    return GS_OK;
}

}  // namespace

extern "C" {

int gs_abi_version(void) { return GS_ABI_VERSION; }

const char *gs_last_error(const gs_ctx *ctx) {
    if (!ctx) return "no context";
    return ctx->err.c_str();
}

gs_status gs_create(gs_ctx **out, int device) {
    if (!out) return GS_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GS_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return GS_ERR_INVALID_ARG;
    gs_ctx *ctx = new gs_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
    // the device's scratch array, with the device current (looked up once per device)
    if (e == hipSuccess) e = gs_scratch_address(device, &ctx->scratch, &ctx->scratch_bytes);
    if (e != hipSuccess) {
        gs_destroy(ctx);
        return GS_ERR_HIP;
    }
    ctx->n_cu = prop.multiProcessorCount;
    ctx->arch = prop.gcnArchName;
    for (char &ch : ctx->arch)
        if (ch == ' ') ch = '_';
    *out = ctx;
    return GS_OK;
}

void gs_destroy(gs_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_words) (void)hipFree(ctx->d_words);
    if (ctx->d_events) (void)hipFree(ctx->d_events);
    if (ctx->d_meta) (void)hipFree(ctx->d_meta);
    if (ctx->d_order) (void)hipFree(ctx->d_order);
    if (ctx->d_order_counts) (void)hipFree(ctx->d_order_counts);
    if (ctx->d_border) (void)hipFree(ctx->d_border);
    if (ctx->d_spill) (void)hipFree(ctx->d_spill);
    if (ctx->d_bcounts) (void)hipFree(ctx->d_bcounts);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    for (int i = 0; i < 3; ++i) {
        if (ctx->side[i]) (void)hipStreamDestroy(ctx->side[i]);
        if (ctx->join[i]) (void)hipEventDestroy(ctx->join[i]);
    }
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->d_params) (void)hipFree(ctx->d_params);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

gs_status gs_default_params(gs_params *p, int32_t model) {
    if (!p || model < GS_MODEL_1D_FP || model > GS_MODEL_2D_GELIOSPHERE) return GS_ERR_INVALID_ARG;
    std::memset(p, 0, sizeof *p);
    const bool oneD = model == GS_MODEL_1D_FP || model == GS_MODEL_1D_BP;
    p->model = model;
    p->injection = GS_INJ_LOG_UNIFORM;
    p->n_bins = 30;
    p->n_weights = 1;
    p->weights[0] = GS_W_POWERLAW;
    p->grid_n = 10000;
    p->K0 = 5e22 * 4.4683705e-27;
    p->V = 400.0 * 6.68458712e-9;
    p->dt = oneD ? 500.0 : 50.0;
    p->r_inj = 1.0;
    p->theta_inj = 90.0 * kPi / 180.0;
    p->tilt = 10.0 * kPi / 180.0;
    p->polarity = 1.0;
    p->ratio = model == GS_MODEL_2D_SOLARPROP ? 0.02 : 0.2;
    p->delta0 = 8.7e-5;
    p->T_min = 0.01;
    p->T_max = 100.0;
    p->n_particles = 1000000;
    p->seed = 1;
    for (int k = 0; k <= 30; ++k) p->bin_edges[k] = 0.01 * std::pow(10.0, 4.0 * k / 30.0);
    return GS_OK;
}

gs_status gs_validate_params(const gs_params *p, char *msg, uint32_t msg_len) {
    std::string m;
    const bool ok = check_params(p, m);
    if (msg && msg_len) {
        std::snprintf(msg, msg_len, "%s", ok ? "ok" : m.c_str());
    }
    return ok ? GS_OK : GS_ERR_INVALID_ARG;
}

gs_status gs_set_launch(gs_ctx *ctx, int32_t model, uint32_t block_threads, uint32_t blocks_per_cu) {
    if (!ctx || model < 0 || model > 3) return GS_ERR_INVALID_ARG;
    for (int k = 0; k < 2; ++k) {
        gs_launch_shape sh = ctx->shape[k][model];
        sh.block_threads = block_threads;
        sh.blocks_per_cu = blocks_per_cu;
        const gs_status st = gs_set_launch_shape(ctx, model, k, &sh);
        if (st != GS_OK) return st;
    }
    return GS_OK;
}

gs_status gs_set_launch_shape(gs_ctx *ctx, int32_t model, int32_t kernel, const gs_launch_shape *shape) {
    if (!ctx || model < 0 || model > 3 || (kernel != GS_KERNEL_PLAIN && kernel != GS_KERNEL_BATCH) || !shape)
        return GS_ERR_INVALID_ARG;
    if (shape->block_threads != 0 && !gs_traj_block_supported(shape->block_threads))
        return fail(ctx, GS_ERR_UNSUPPORTED, "block_threads must be 0, 128, 256 or 512");
    if (shape->blocks_per_cu > 64) return fail(ctx, GS_ERR_INVALID_ARG, "blocks_per_cu above 64");
    if (shape->chunk > 4096) return fail(ctx, GS_ERR_INVALID_ARG, "chunk above 4096");
    ctx->shape[kernel][model] = *shape;
    ctx->tuned_rate[kernel][model] = 0.0;
    return GS_OK;
}

gs_status gs_get_launch_shape(const gs_ctx *ctx, int32_t model, int32_t kernel, gs_launch_shape *shape) {
    if (!ctx || model < 0 || model > 3 || (kernel != GS_KERNEL_PLAIN && kernel != GS_KERNEL_BATCH) || !shape)
        return GS_ERR_INVALID_ARG;
    HIP_TRY(const_cast<gs_ctx *>(ctx), hipSetDevice(ctx->device));
    const Resolved R = resolve_shape(ctx, model, kernel == GS_KERNEL_BATCH);
    shape->block_threads = R.block;
    shape->blocks_per_cu = R.bpcu;
    shape->chunk = R.chunk;
    return GS_OK;
}

gs_status gs_set_order(gs_ctx *ctx, int32_t mode) {
    if (!ctx) return GS_ERR_INVALID_ARG;
    if (mode != GS_ORDER_ID && mode != GS_ORDER_ENERGY)
        return fail(ctx, GS_ERR_INVALID_ARG, "unknown order mode %d", (int)mode);
    ctx->order_mode = mode;
    return GS_OK;
}

gs_status gs_autotune_shape(gs_ctx *ctx, const gs_params *pilots, uint32_t n_pilots, int32_t kernel,
                            uint32_t repeats, gs_launch_shape *best_shape, double *best_steps_per_s) {
    if (!ctx || !pilots || n_pilots == 0) return GS_ERR_INVALID_ARG;
    if (kernel != GS_KERNEL_PLAIN && kernel != GS_KERNEL_BATCH) return fail(ctx, GS_ERR_INVALID_ARG, "unknown kernel");
    if (kernel == GS_KERNEL_PLAIN && n_pilots != 1) return fail(ctx, GS_ERR_INVALID_ARG, "plain tuning takes one pilot");
    const int model = pilots[0].model;
    for (uint32_t i = 0; i < n_pilots; ++i) {
        std::string msg;
        if (!check_params(&pilots[i], msg)) return fail(ctx, GS_ERR_INVALID_ARG, "invalid pilot %u: %s", i, msg.c_str());
        if (pilots[i].model != model) return fail(ctx, GS_ERR_INVALID_ARG, "pilots of more than one model");
    }
    if (repeats == 0) repeats = 1;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const bool batch = kernel == GS_KERNEL_BATCH;
    const gs_launch_shape saved = ctx->shape[kernel][model];
    std::vector<gs_spectrum> outs(n_pilots);
    // best rate (particle-steps/s of trajectory kernel time) of one shape
    auto measure = [&](const gs_launch_shape &sh, double *rate) -> gs_status {
        ctx->shape[kernel][model] = sh;
        double best = 0.0;
        for (uint32_t r = 0; r < repeats; ++r) {
            const gs_status st = batch ? gs_run_batch(ctx, pilots, n_pilots, outs.data())
                                       : gs_run_plain(ctx, pilots, outs.data(), nullptr, 0, nullptr);
            if (st != GS_OK) return st;
            uint64_t steps = 0;
            for (const gs_spectrum &o : outs) steps += o.total_steps;
            const double v = ctx->stats.kernel_ms > 0 ? (double)steps / (ctx->stats.kernel_ms * 1e-3) : 0.0;
            best = std::max(best, v);
        }
        *rate = best;
        return GS_OK;
    };
    gs_launch_shape best{256, 0, (uint32_t)gs::kChunk};
    double best_rate = -1.0;
    // GS_AUTOTUNE_LOG set: one line per measured shape on stderr
    const char *log_env = std::getenv("GS_AUTOTUNE_LOG");
    const bool log = log_env && *log_env && std::strcmp(log_env, "0") != 0;
    auto consider = [&](const gs_launch_shape &sh) -> gs_status {
        double v = 0.0;
        const gs_status st = measure(sh, &v);
        if (st == GS_OK && v > best_rate) { best_rate = v; best = sh; }
        if (log)
            std::fprintf(stderr, "gs_autotune: model %d %s: %u threads, %u blocks/CU, grab %u: %.4g particle-steps/s\n",
                         model, batch ? "batch" : "plain", sh.block_threads, sh.blocks_per_cu, sh.chunk, v);
        return st;
    };
    gs_status st = GS_OK;
    // 1. blocks per CU at 256 threads and the default grab
    int occ = 0;
    HIP_TRY(ctx, gs_traj_occupancy(model, batch, 256, &occ));
    for (int b = 1; b <= std::max(occ, 1) && st == GS_OK; ++b) st = consider({256, (uint32_t)b, (uint32_t)gs::kChunk});
    // 2. the other workgroup sizes at about the same waves per CU, and at their occupancy
    const unsigned waves = best.blocks_per_cu * 4;
    for (unsigned blk : {128u, 512u}) {
        int o = 0;
        HIP_TRY(ctx, gs_traj_occupancy(model, batch, blk, &o));
        const int eq = (int)((waves * 64 + blk / 2) / blk);
        std::vector<int> cand = {eq - 1, eq, eq + 1, o};
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (int b : cand)
            if (b >= 1 && b <= o && st == GS_OK) st = consider({blk, (uint32_t)b, (uint32_t)gs::kChunk});
    }
    // 3. ids per queue grab
    const gs_launch_shape base = best;
    for (uint32_t c : {16u, 32u, 128u, 256u})
        if (st == GS_OK) st = consider({base.block_threads, base.blocks_per_cu, c});
    if (st != GS_OK) {
        ctx->shape[kernel][model] = saved;
        return st;
    }
    ctx->shape[kernel][model] = best;
    ctx->tuned_rate[kernel][model] = best_rate;
    if (best_shape) *best_shape = best;
    if (best_steps_per_s) *best_steps_per_s = best_rate;
    return GS_OK;
}

gs_status gs_autotune(gs_ctx *ctx, const gs_params *pilot, uint32_t repeats, uint32_t *best_bpcu,
                      double *best_steps_per_s) {
    gs_launch_shape sh;
    const gs_status st = gs_autotune_shape(ctx, pilot, 1, GS_KERNEL_PLAIN, repeats, &sh, best_steps_per_s);
    if (st == GS_OK && best_bpcu) *best_bpcu = sh.blocks_per_cu;
    return st;
}

// Cache lines: "<arch> <CUs> <plain|batch> <model> <block_threads> <blocks_per_cu> <chunk> <steps/s>".
gs_status gs_autotune_save(const gs_ctx *ctx, const char *path) {
    if (!ctx || !path) return GS_ERR_INVALID_ARG;
    std::vector<std::string> lines;
    if (FILE *in = std::fopen(path, "r")) {   // keep other devices' and untuned entries
        char buf[512];
        while (std::fgets(buf, sizeof buf, in)) {
            char arch[256], kind[16];
            int cus, model;
            if (std::sscanf(buf, "%255s %d %15s %d", arch, &cus, kind, &model) == 4 && ctx->arch == arch &&
                cus == ctx->n_cu && model >= 0 && model < 4 &&
                (std::strcmp(kind, "batch") == 0 || std::strcmp(kind, "plain") == 0)) {
                const int k = std::strcmp(kind, "batch") == 0 ? 1 : 0;
                if (ctx->tuned_rate[k][model] > 0) continue;   // replaced below
            }
            lines.emplace_back(buf);
            if (lines.back().empty() || lines.back().back() != '\n') lines.back() += '\n';
        }
        std::fclose(in);
    }
    FILE *out = std::fopen(path, "w");
    if (!out) return fail(const_cast<gs_ctx *>(ctx), GS_ERR_IO, "cannot write %s", path);
    if (lines.empty()) std::fputs("# arch CUs kernel model block_threads blocks_per_cu chunk particle_steps_per_s\n", out);
    for (const std::string &l : lines) std::fputs(l.c_str(), out);
    for (int k = 0; k < 2; ++k)
        for (int m = 0; m < 4; ++m)
            if (ctx->tuned_rate[k][m] > 0) {
                const gs_launch_shape &sh = ctx->shape[k][m];
                std::fprintf(out, "%s %d %s %d %u %u %u %.6g\n", ctx->arch.c_str(), ctx->n_cu, k ? "batch" : "plain",
                             m, sh.block_threads, sh.blocks_per_cu, sh.chunk, ctx->tuned_rate[k][m]);
            }
    std::fclose(out);
    return GS_OK;
}

gs_status gs_autotune_load(gs_ctx *ctx, const char *path, uint32_t *n_applied) {
    if (!ctx || !path) return GS_ERR_INVALID_ARG;
    if (n_applied) *n_applied = 0;
    FILE *in = std::fopen(path, "r");
    if (!in) return fail(ctx, GS_ERR_IO, "cannot read %s", path);
    char buf[512];
    uint32_t n = 0;
    while (std::fgets(buf, sizeof buf, in)) {
        char arch[256], kind[16];
        int cus, model;
        unsigned blk, bpcu, chunk;
        double rate;
        if (std::sscanf(buf, "%255s %d %15s %d %u %u %u %lf", arch, &cus, kind, &model, &blk, &bpcu, &chunk,
                        &rate) != 8)
            continue;
        if (ctx->arch != arch || cus != ctx->n_cu || model < 0 || model > 3) continue;
        const bool is_batch = std::strcmp(kind, "batch") == 0;
        if (!is_batch && std::strcmp(kind, "plain") != 0) continue;   // not a kernel this library has
        const int k = is_batch ? 1 : 0;
        const gs_launch_shape sh{blk, bpcu, chunk};
        if (gs_set_launch_shape(ctx, model, k, &sh) != GS_OK) continue;
        ctx->tuned_rate[k][model] = rate;
        ++n;
    }
    std::fclose(in);
    if (n_applied) *n_applied = n;
    return GS_OK;
}

gs_status gs_spectrum_merge(const gs_spectrum *parts, uint32_t n, gs_spectrum *out) {
    if (!parts || !out || n == 0) return GS_ERR_INVALID_ARG;
    const uint32_t nb = parts[0].n_bins, nw = parts[0].n_weights;
    if (nb > GS_MAX_BINS || nw > GS_MAX_WEIGHTS) return GS_ERR_INVALID_ARG;
    for (uint32_t j = 1; j < n; ++j)
        if (parts[j].n_bins != nb || parts[j].n_weights != nw) return GS_ERR_INVALID_ARG;
    gs_spectrum *m = new gs_spectrum;   // out may alias a part
    std::memset(m, 0, sizeof *m);
    m->n_bins = nb;
    m->n_weights = nw;
    for (uint32_t j = 0; j < n; ++j) {
        const gs_spectrum &p = parts[j];
        for (uint32_t b = 0; b < nb; ++b) m->count[b] += p.count[b];
        m->n_particles += p.n_particles;
        m->n_events += p.n_events;
        m->n_outside += p.n_outside;
        m->total_steps += p.total_steps;
        m->n_nan += p.n_nan;
        m->n_step_cap += p.n_step_cap;
        m->n_acc_inexact += p.n_acc_inexact;
        m->n_acc_overflow += p.n_acc_overflow;
        for (uint32_t k = 0; k < nw; ++k)
            for (int s = 0; s < 2; ++s)
                for (uint32_t b = 0; b < nb; ++b) {
                    uint32_t *L = m->acc[k][s][b];
                    const uint32_t *A = p.acc[k][s][b];
                    uint64_t carry = 0;
                    for (int i = 0; i < GS_ACC_LIMBS; ++i) {
                        const uint64_t v = (uint64_t)L[i] + A[i] + carry;
                        L[i] = (uint32_t)v;
                        carry = v >> 32;
                    }
                    if (carry) ++m->n_acc_overflow;
                }
    }
    for (uint32_t k = 0; k < nw; ++k)
        for (uint32_t b = 0; b < nb; ++b) {
            m->sum_w[k][b] = limbs_to_double(m->acc[k][0][b]);
            m->sum_w2[k][b] = limbs_to_double(m->acc[k][1][b]);
        }
    const bool overflow = m->n_acc_overflow != 0;
    *out = *m;
    delete m;
    return overflow ? GS_ERR_OVERFLOW : GS_OK;
}

gs_status gs_last_run_stats(const gs_ctx *ctx, gs_run_stats *st) {
    if (!ctx || !st) return GS_ERR_INVALID_ARG;
    *st = ctx->stats;
    return GS_OK;
}

gs_status gs_run_plain(gs_ctx *ctx, const gs_params *p, gs_spectrum *out, gs_exit_record *recs,
                       uint64_t rec_cap, uint64_t *n_recs) {
    if (!ctx) return GS_ERR_INVALID_ARG;
    if (!out) return fail(ctx, GS_ERR_INVALID_ARG, "out is NULL");
    std::string msg;
    if (!check_params(p, msg)) return fail(ctx, GS_ERR_INVALID_ARG, "invalid params: %s", msg.c_str());
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    gs_status st = ensure_sims(ctx, 1);
    if (st == GS_OK) st = ensure_params(ctx, 1);
    if (st == GS_OK) st = ensure_meta(ctx, 64);
    if (st != GS_OK) return st;
    PlainBufs B;
    B.words = ctx->d_words;
    B.params = ctx->d_params;
    B.meta = ctx->d_meta;
    B.own = true;
    B.events = nullptr;
    B.cap_events = 0;
    B.order_counts = nullptr;
    B.order = nullptr;
    return run_plain_core(ctx, *p, out, recs, rec_cap, n_recs, B);
}

gs_status gs_run_plain_noalloc(gs_ctx *ctx, const gs_params *p, gs_spectrum *out) {
    if (!ctx) return GS_ERR_INVALID_ARG;
    if (!out) return fail(ctx, GS_ERR_INVALID_ARG, "out is NULL");
    if (p && (p->n_bins > GS_MAX_BINS || p->n_weights > GS_MAX_WEIGHTS))
        return fail(ctx, GS_ERR_INVALID_ARG, "histogram of %u bins x %u weights exceeds the scratch maximum of %u x %u",
                    (unsigned)p->n_bins, (unsigned)p->n_weights, (unsigned)GS_MAX_BINS, (unsigned)GS_MAX_WEIGHTS);
    std::string msg;
    if (!check_params(p, msg)) return fail(ctx, GS_ERR_INVALID_ARG, "invalid params: %s", msg.c_str());
    if (!ctx->scratch || ctx->scratch_bytes < gs::kScratchWords * sizeof(uint64_t))
        return fail(ctx, GS_ERR_HIP, "no scratch array on device %d", ctx->device);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    uint64_t *base = (uint64_t *)ctx->scratch;   // the device address, never the symbol
    PlainBufs B;
    B.words = base;
    B.params = (DevParams *)(base + kScratchParamsAt);
    B.meta = base + kScratchMetaAt;
    B.own = false;
    B.events = (gs::DevEvent *)(base + kScratchEventsAt);
    B.cap_events = gs::kScratchEvents;
    B.order_counts = (uint32_t *)(base + kScratchCountsAt);
    B.order = (uint32_t *)(base + kScratchPermAt);
    return run_plain_core(ctx, *p, out, nullptr, 0, nullptr, B);
}

gs_status gs_debug_math(gs_ctx *ctx, int32_t fn, const double *x, double *y, uint64_t n) {
    if (!ctx || (!x && fn != GS_DBG_NORMALS) || !y) return GS_ERR_INVALID_ARG;
    if (fn < GS_DBG_RCP || fn > GS_DBG_RSQRT4) return fail(ctx, GS_ERR_INVALID_ARG, "unknown fn");
    if (!x) return GS_ERR_INVALID_ARG;
    if (n == 0) return GS_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const bool two = fn == GS_DBG_ATAN_FRAC || fn == GS_DBG_SHEET_F || fn == GS_DBG_SHEET_G;
    const uint64_t nin = two ? 2 * n : n;   // two-argument forms: 2n inputs
    double *d = nullptr;
    HIP_TRY(ctx, hipMalloc((void **)&d, (nin + n) * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, x, nin * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = gs_launch_debug(fn, d, d + nin, n, 0, 0, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(y, d + nin, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, GS_ERR_HIP, "gs_debug_math: %s", hipGetErrorString(e));
    return GS_OK;
}

gs_status gs_debug_normals(gs_ctx *ctx, uint64_t seed, uint64_t id, uint64_t n, double *out) {
    if (!ctx || !out) return GS_ERR_INVALID_ARG;
    if (n == 0) return GS_OK;
    if (n > (1ull << 32)) return fail(ctx, GS_ERR_INVALID_ARG, "n above 2^32");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    double *d = nullptr;
    HIP_TRY(ctx, hipMalloc((void **)&d, n * sizeof(double)));
    hipError_t e = gs_launch_debug(GS_DBG_NORMALS, nullptr, d, n, seed, id, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, GS_ERR_HIP, "gs_debug_normals: %s", hipGetErrorString(e));
    return GS_OK;
}

// The device kernels' Box-Muller tables, for the host evaluation.
alignas(16) static const double kBmTrigHost[gs::fm::kBmTrigWords] = GS_BM_TRIG_TABLE;
alignas(16) static const double kBmLogHost[gs::fm::kBmLogWords] = GS_BM_LOG_TABLE;

gs_status gs_host_math(int32_t fn, const double *x, double *y, uint64_t n) {
    if (!x || !y) return GS_ERR_INVALID_ARG;
    namespace fm = gs::fm;
    using H = fm::HostOps;
    for (uint64_t i = 0; i < n; ++i) {
        double s, c;
        switch (fn) {
        case GS_DBG_LOG: y[i] = fm::log_e<H>(x[i], 0); break;
        case GS_DBG_EXP: y[i] = fm::exp_e<H>(x[i]); break;
        case GS_DBG_SIN_TURN: fm::sincos_turn<H>((uint32_t)x[i], kBmTrigHost, s, c); y[i] = s; break;
        case GS_DBG_COS_TURN: fm::sincos_turn<H>((uint32_t)x[i], kBmTrigHost, s, c); y[i] = c; break;
        case GS_DBG_LOG_U32: y[i] = fm::log_u32<H>((uint32_t)x[i], kBmLogHost); break;
        case GS_DBG_SIN: fm::sincos_tab<H>(x[i], kBmTrigHost, s, c); y[i] = s; break;
        case GS_DBG_COS: fm::sincos_tab<H>(x[i], kBmTrigHost, s, c); y[i] = c; break;
        case GS_DBG_ATAN: y[i] = fm::atan_e<H>(x[i]); break;
        case GS_DBG_ATAN_FRAC: y[i] = fm::atan_frac<H>(x[i], x[n + i], fm::kAtanRed); break;
        case GS_DBG_SHEET_F: fm::sheet_drift<H>(x[i], x[n + i], fm::kAtanRed, s, c); y[i] = s; break;
        case GS_DBG_SHEET_G: fm::sheet_drift<H>(x[i], x[n + i], fm::kAtanRed, s, c); y[i] = c; break;
        default: return GS_ERR_UNSUPPORTED;   // rcp / rsqrt / sqrt / rsqrt4 seeds exist on the device only
        }
    }
    return GS_OK;
}

gs_status gs_set_batch_budget(gs_ctx *ctx, uint64_t max_events) {
    if (!ctx) return GS_ERR_INVALID_ARG;
    if (max_events < (1u << 16)) return fail(ctx, GS_ERR_INVALID_ARG, "budget below 65536 events");
    ctx->batch_event_budget = max_events;
    return GS_OK;
}

gs_status gs_set_plain_limits(gs_ctx *ctx, uint64_t max_chunk, uint64_t max_events) {
    if (!ctx) return GS_ERR_INVALID_ARG;
    if (max_chunk > kMaxChunk) return fail(ctx, GS_ERR_INVALID_ARG, "max_chunk above 2^27");
    ctx->plain_chunk = max_chunk ? max_chunk : kMaxChunk;
    ctx->plain_event_cap = max_events;
    return GS_OK;
}

gs_status gs_run_batch_records(gs_ctx *ctx, const gs_params *sims, uint32_t n_sims, gs_spectrum *outs,
                               gs_exit_record *const *recs, const uint64_t *rec_caps, uint64_t *n_recs) {
    if (!ctx) return GS_ERR_INVALID_ARG;
    if ((!sims || !outs) && n_sims) return fail(ctx, GS_ERR_INVALID_ARG, "NULL argument");
    for (uint32_t i = 0; i < n_sims; ++i) {
        std::string msg;
        if (!check_params(&sims[i], msg))
            return fail(ctx, GS_ERR_INVALID_ARG, "invalid params of simulation %u: %s", i, msg.c_str());
    }
    if (recs && !rec_caps) return fail(ctx, GS_ERR_INVALID_ARG, "records without capacities");
    const BatchRecs R{recs, rec_caps, n_recs};
    const auto t_start = std::chrono::steady_clock::now();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    gs_run_stats acc{};
    acc.block_threads = 0;
    // Simulations too large for one event region run on their own.
    std::vector<uint32_t> small;
    for (uint32_t i = 0; i < n_sims; ++i) {
        const gs_params &p = sims[i];
        if (p.n_particles > kMaxChunk || event_room(p.model, p.n_particles) > ctx->batch_event_budget / 2) {
            gs_status st = gs_run_plain(ctx, &p, &outs[i], R.want(i) ? recs[i] : nullptr, R.want(i) ? rec_caps[i] : 0,
                                        n_recs ? &n_recs[i] : nullptr);
            if (st != GS_OK) return st;
            acc.kernel_ms += ctx->stats.kernel_ms;
            acc.lane_cycles += ctx->stats.lane_cycles;
            acc.active_cycles += ctx->stats.active_cycles;
        } else {
            small.push_back(i);
        }
    }
    // Largest estimated work first, then cut into sub-batches that fit the
    // event budget (one persistent launch group each).
    std::stable_sort(small.begin(), small.end(), [&](uint32_t a, uint32_t b) {
        return est_cost(sims[a]) * (double)sims[a].n_particles > est_cost(sims[b]) * (double)sims[b].n_particles;
    });
    size_t i0 = 0;
    while (i0 < small.size()) {
        uint64_t room = 0;
        size_t i1 = i0;
        while (i1 < small.size() && i1 - i0 < kMaxBatchSims) {
            const uint64_t r = event_room(sims[small[i1]].model, sims[small[i1]].n_particles);
            if (i1 > i0 && room + r > ctx->batch_event_budget) break;
            room += r;
            ++i1;
        }
        std::vector<uint32_t> part(small.begin() + i0, small.begin() + i1);
        std::vector<uint32_t> redo;
        gs_status st = run_sub_batch(ctx, sims, part, outs, redo, acc, R);
        if (st != GS_OK) return st;
        for (uint32_t i : redo) {   // event region overflow (1D F-p): run alone
            st = gs_run_plain(ctx, &sims[i], &outs[i], R.want(i) ? recs[i] : nullptr, R.want(i) ? rec_caps[i] : 0,
                              n_recs ? &n_recs[i] : nullptr);
            if (st != GS_OK) return st;
            acc.kernel_ms += ctx->stats.kernel_ms;
            acc.lane_cycles += ctx->stats.lane_cycles;
            acc.active_cycles += ctx->stats.active_cycles;
        }
        i0 = i1;
    }
    acc.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    ctx->stats = acc;
    for (uint32_t i = 0; i < n_sims; ++i)
        if (outs[i].n_acc_overflow) return fail(ctx, GS_ERR_OVERFLOW, "weight accumulator overflow (simulation %u)", i);
    return GS_OK;
}

gs_status gs_run_batch(gs_ctx *ctx, const gs_params *sims, uint32_t n_sims, gs_spectrum *outs) {
    return gs_run_batch_records(ctx, sims, n_sims, outs, nullptr, nullptr, nullptr);
}

}  // extern "C"
