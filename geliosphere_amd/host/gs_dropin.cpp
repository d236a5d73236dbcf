// gs_dropin.cpp - the four GPU model adapters and their shared host logic.
#include "gs_dropin.hpp"

#include <dirent.h>
#include <errno.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>

#include "GeliosphereGpuModel.hpp"
#include "OneDimensionBpGpuModel.hpp"
#include "OneDimensionFpGpuModel.hpp"
#include "SolarPropLikeGpuModel.hpp"

namespace gsdrop {

namespace {

// Physical constants of the reference (Algorithm/include/ResultConstants.hpp:19-29).
const double kM0 = 1.67261e-27;
const double kQ = 1.60219e-19;
const double kC = 2.99793e8;
const double kT0w = kM0 * kC * kC;

// Momentum [kg m/s] of a proton of kinetic energy T [GeV]
// (OneDimensionFpModel.cu:35-37 in double).
double momentum_si(double T) {
    const double Tkinw = T * 1e9 * kQ;
    const double Rig = std::sqrt(Tkinw * (Tkinw + (2.0 * kT0w))) / kQ;
    return Rig * kQ / kC;
}

// The per-line weight the reference GPU models print in log.dat:
// ((m0^2 c^4 + p^2 c^2)^-1.85 / p) / 1e45 (OneDimensionFpModel.cu:42-43).
double log_weight(double p) {
    const double w = (kM0 * kM0 * kC * kC * kC * kC) + (p * p * kC * kC);
    return (std::pow(w, -1.85) / p) / 1e45;
}

void log_info(const std::string &s) { std::fprintf(stderr, "[info] %s\n", s.c_str()); }
void log_error(const std::string &s) { std::fprintf(stderr, "[error] %s\n", s.c_str()); }

uint64_t clock_seed() {
    return (uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count() ^
           ((uint64_t)std::clock() << 32);
}

int mkdir_and_chdir(const std::string &name) {
    DIR *dir = opendir(name.c_str());
    if (dir) {
        closedir(dir);
    } else if (errno == ENOENT) {
        if (mkdir(name.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH) != 0) {
            std::fprintf(stderr, "ERROR: unable to mkdir %s: %s\n", name.c_str(), std::strerror(errno));
            return -1;
        }
    } else {
        std::fprintf(stderr, "ERROR: unable to open directory %s\n", name.c_str());
        return -1;
    }
    if (chdir(name.c_str()) != 0) {
        std::fprintf(stderr, "ERROR: unable to change directory to %s: %s\n", name.c_str(), std::strerror(errno));
        return -1;
    }
    return 1;
}

}  // namespace

int model_from_name(const std::string &name) {
    if (name == "1D Fp") return GS_MODEL_1D_FP;
    if (name == "1D Bp") return GS_MODEL_1D_BP;
    if (name == "2D SolarProp-like") return GS_MODEL_2D_SOLARPROP;
    if (name == "2D Geliosphere") return GS_MODEL_2D_GELIOSPHERE;
    return -1;
}

const char *model_name(int model) {
    switch (model) {
    case GS_MODEL_1D_FP: return "1D Fp";
    case GS_MODEL_1D_BP: return "1D Bp";
    case GS_MODEL_2D_SOLARPROP: return "2D SolarProp-like";
    case GS_MODEL_2D_GELIOSPHERE: return "2D Geliosphere";
    default: return "?";
    }
}

const char *model_directory(int model) {
    switch (model) {
    case GS_MODEL_1D_FP: return "1DFP";
    case GS_MODEL_1D_BP: return "1DBP";
    case GS_MODEL_2D_SOLARPROP: return "SOLARPROPLike";
    case GS_MODEL_2D_GELIOSPHERE: return "Geliosphere2D";
    default: return "unknown";
    }
}

bool config_from_carrier(ParamsCarrier *pc, int model, RunConfig *cfg, std::string *err) {
    if (model < GS_MODEL_1D_FP || model > GS_MODEL_2D_GELIOSPHERE) {
        *err = "unknown model";
        return false;
    }
    gs_params &p = cfg->params;
    gs_default_params(&p, model);
    const bool two_d = model == GS_MODEL_2D_SOLARPROP || model == GS_MODEL_2D_GELIOSPHERE;
    // Values the reference reads as floats stay floats until they reach the
    // double parameter block, so a run sees the same numbers as the
    // reference's CPU path (Constants.hpp:198-218).
    const float dt = pc->getFloat("dt", pc->getFloat("dt_default", -1.0f));
    p.dt = dt != -1.0f ? dt : 5.0f;   // CosmicConstants.cu:15 when nothing sets it
    const float K = pc->getFloat("K0", pc->getFloat("K0_default", -1.0f));
    const bool user_K = pc->getInt("K0_entered_by_user", 0) != 0;
    float K0 = 0.000222f;             // CosmicConstants.cu:16
    if (K != -1.0f && !user_K && model == GS_MODEL_2D_SOLARPROP) K0 = K;                      // :97-101
    if (K != -1.0f && !user_K && model == GS_MODEL_2D_GELIOSPHERE) K0 = K * pc->getFloat("K0_ratio", 5.0f);  // :105-110
    if (K != -1.0f && user_K) K0 = K;                                                         // :69-73
    p.K0 = K0;
    const bool back = model == GS_MODEL_1D_BP;
    float V = back ? pc->getFloat("V", pc->getFloat("V_default", 1.0f)) * (-1.0f)
                   : pc->getFloat("V", pc->getFloat("V_default", -1.0f));
    if (V == -1.0f) V = back ? (float)(2.66667e-6 * (-1.0f)) : (float)2.66667e-6;   // :74-83
    p.V = back ? -(double)V : (double)V;
    p.r_inj = pc->getFloat("r_injection", 1.0f);
    const float theta = pc->getFloat("theta_injection", 90.0f) * 3.1415926535f / 180.0f;
    cfg->theta_inj_f = theta;
    p.theta_inj = theta;
    p.polarity = (float)pc->getInt("polarity", 1);
    const float tilt = pc->getFloat("tilt_angle", pc->getFloat("default_tilt_angle", 0.1f)) * 3.1415926535f / 180.0f;
    if (model == GS_MODEL_2D_SOLARPROP) {
        p.ratio = pc->getFloat("solarPropRatio", 0.02f);
        if (pc->getInt("gs_solarprop_gpu_compat", 0)) {
            // The reference GPU kernel never sets the tilt for this model and
            // drifts theta with polarity +1 (CosmicConstants.cu:28,37).
            p.tilt = 5.75f * 3.1415926535f / 180.0f;
            p.flags |= GS_FLAG_SOLARPROP_GPU_DRIFT;
        } else {
            p.tilt = tilt;   // Constants.hpp:160
        }
    }
    if (model == GS_MODEL_2D_GELIOSPHERE) {
        p.ratio = pc->getFloat("geliosphereRatio", 0.2f);
        p.delta0 = pc->getFloat("C_delta", 8.7e-5f);
        p.tilt = tilt;
    }
    // Injection energies (CosmicUtils.cu:17-40): 1D models always, 2D models
    // with use_uniform_injection, take the 1e-4 GeV grid up to
    // uniform_energy_injection_maximum; 2D models otherwise cycle through
    // the 30 SolarProp energies.
    const float inj_max = pc->getFloat("uniform_energy_injection_maximum", 101.0f);
    if (!two_d || pc->getInt("use_uniform_injection", 0)) {
        p.injection = GS_INJ_GRID;
        p.grid_n = (uint32_t)std::lround((double)inj_max * 10000.0);
        p.T_min = 1e-4;
        p.T_max = 1e-4 + p.grid_n * 1e-4;
    } else {
        p.injection = GS_INJ_SOLARPROP_BINS;
    }
    // Particles: "millions" rounded up to whole launches of 4096 x 256, as
    // the reference GPU models run them (OneDimensionBpModel.cu:170).
    const int exact = pc->getInt("gs_particles", -1);
    if (exact >= 0) {
        p.n_particles = (uint64_t)exact;
    } else {
        const double millions = pc->getInt("millions", 1);
        const double per_launch = 4096.0 * 256.0;
        p.n_particles = (uint64_t)std::ceil(millions * 1000000.0 / per_launch) * (uint64_t)per_launch;
    }
    // Spectrum: 30 log bins, 0.01-100 GeV, on the energy the Results pass
    // bins by; weights of the Results pass of the model.
    p.n_bins = 30;
    for (int k = 0; k <= 30; ++k) p.bin_edges[k] = 0.01 * std::pow(10.0, 4.0 * k / 30.0);
    if (two_d) {
        p.n_weights = 3;
        p.weights[0] = GS_W_WEBER;
        p.weights[1] = GS_W_USOSKIN;
        p.weights[2] = GS_W_BURGER;
    } else {
        p.n_weights = 2;
        p.weights[0] = GS_W_YAMADA;
        p.weights[1] = GS_W_POWERLAW;
    }
    const std::string seed = pc->getString("gs_seed", "");
    cfg->seed = seed.empty() ? clock_seed() : std::strtoull(seed.c_str(), nullptr, 10);
    p.seed = cfg->seed;
    cfg->write_log = pc->getInt("gs_write_log", 1) != 0;
    cfg->device = pc->getInt("gs_device", 0);
    char msg[256];
    if (gs_validate_params(&p, msg, sizeof msg) != GS_OK) {
        *err = msg;
        return false;
    }
    return true;
}

bool create_directory(const std::string &model_dir, const std::string &destination) {
    if (mkdir_and_chdir("results") != 1) return false;
    if (mkdir_and_chdir(model_dir) != 1) return false;
    if (mkdir_and_chdir(destination) != 1) return false;
    return true;
}

std::string directory_name(ParamsCarrier *pc) {
    const time_t clk = time(nullptr);
    std::string name = ctime(&clk);
    if (!name.empty() && name.back() == '\n') name.pop_back();
    std::replace(name.begin(), name.end(), ' ', '_');
    name += "_dt=" + std::to_string(pc->getFloat("dt", pc->getFloat("dt_default", 50.0f))) +
            "K0=" + std::to_string(pc->getFloat("K0", pc->getFloat("K0_default", 0.000222f))) +
            "V=" + std::to_string(pc->getFloat("V", pc->getFloat("V_default", 2.66667e-6f)));
    return name;
}

std::string simulation_report(ParamsCarrier *pc, const RunConfig &cfg, const char *engine) {
    std::string out;
    char line[256];
    auto add = [&](const char *fmt, auto... v) {
        std::snprintf(line, sizeof line, fmt, v...);
        out += line;
    };
    const std::string model = pc->getString("model", "1D Fp");
    add("Model: %s\n", model.c_str());
    out += "---------------------\n";
    if (pc->getInt("month_option", -1) != -1) {
        add("Selected month: %d\n", pc->getInt("month_option", -1));
        add("Selected year: %d\n", pc->getInt("year_option", -1));
    }
    if (model == "2D SolarProp-like") add("SolarProp-like model ratio: %g\n", pc->getFloat("solarPropRatio", 0.02f));
    if (model == "2D Geliosphere") {
        add("Geliosphere model ratio: %g\n", pc->getFloat("geliosphereRatio", 0.2f));
        add("C Delta: %g\n", pc->getFloat("C_delta", 8.7e-5f));
        const float tilt = pc->getFloat("tilt_angle", pc->getFloat("default_tilt_angle", 0.1f));
        add("Loaded tilt angle from file: %g\n", tilt);
        add("Calculated tilt angle:: %g\n", tilt * 3.1415926535f / 180.0f);
    }
    add("dt: %g\n", pc->getFloat("dt", pc->getFloat("dt_default", 50.0f)));
    add("K0: %g\n", pc->getFloat("K0", pc->getFloat("K0_default", 0.000222f)));
    add("V: %g\n", pc->getFloat("V", pc->getFloat("V_default", 2.66667e-6f)));
    add("Engine: %s\n", engine);
    add("Seed: %llu\n", (unsigned long long)cfg.seed);
    add("Test particles: %llu\n", (unsigned long long)cfg.params.n_particles);
    return out;
}

void write_simulation_report(ParamsCarrier *pc, const RunConfig &cfg) {
    FILE *f = std::fopen("simulation_log.txt", "w");
    if (!f) return;
    const std::string text = simulation_report(pc, cfg, "gfx950 SDE trajectory engine (FP64, Philox4x32-10)");
    std::fputs(text.c_str(), f);
    std::fclose(f);
}

bool write_log_dat(const char *path, int model, const std::vector<gs_exit_record> &recs, float theta_inj_f) {
    FILE *f = std::fopen(path, "w");
    if (!f) return false;
    std::vector<char> buf(1 << 20);
    std::setvbuf(f, buf.data(), _IOFBF, buf.size());
    for (const gs_exit_record &e : recs) {
        switch (model) {
        case GS_MODEL_1D_BP:
            std::fprintf(f, "%g %g %g %g\n", e.t_exit, e.t_inj, e.r, log_weight(momentum_si(e.t_exit)));
            break;
        case GS_MODEL_1D_FP: {
            const double pinj = momentum_si(e.t_inj);
            std::fprintf(f, " %g  %g  %g  %g %g \n", pinj, momentum_si(e.t_exit), e.r, log_weight(pinj), e.sumac);
            break;
        }
        default:
            std::fprintf(f, "%g %g %g %g %g %g\n", e.t_inj, e.t_exit, e.r, log_weight(momentum_si(e.t_inj)),
                         (double)theta_inj_f, e.theta);
            break;
        }
    }
    std::fclose(f);
    return true;
}

bool write_histogram(const std::string &base, const gs_params &p, const gs_spectrum &s, bool csv) {
    static const char *wname[] = {"powerlaw", "yamada", "weber", "usoskin", "burger"};
    FILE *f = std::fopen((base + (csv ? ".csv" : ".dat")).c_str(), "w");
    if (!f) return false;
    const char *sep = csv ? "," : " ";
    std::fprintf(f, "%sT_lo%sT_hi%scount", csv ? "" : "# ", sep, sep);
    for (uint32_t k = 0; k < s.n_weights; ++k)
        std::fprintf(f, "%ssum_w_%s%ssum_w2_%s%smean_%s%sse_%s", sep, wname[p.weights[k]], sep,
                     wname[p.weights[k]], sep, wname[p.weights[k]], sep, wname[p.weights[k]]);
    std::fprintf(f, "\n");
    for (uint32_t b = 0; b < s.n_bins; ++b) {
        std::fprintf(f, "%.17g%s%.17g%s%llu", p.bin_edges[b], sep, p.bin_edges[b + 1], sep,
                     (unsigned long long)s.count[b]);
        const double n = (double)s.count[b];
        for (uint32_t k = 0; k < s.n_weights; ++k) {
            const double sw = s.sum_w[k][b], sw2 = s.sum_w2[k][b];
            const double mean = n > 0 ? sw / n : 0.0;
            const double var = n > 1 ? std::max(0.0, (sw2 / n - mean * mean)) : 0.0;
            const double se = n > 0 ? std::sqrt(var / n) : 0.0;
            std::fprintf(f, "%s%.17g%s%.17g%s%.17g%s%.17g", sep, sw, sep, sw2, sep, mean, sep, se);
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return true;
}

gs_ctx *context(int device) {
    static std::map<int, gs_ctx *> ctxs;
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second;
    gs_ctx *ctx = nullptr;
    const gs_status st = gs_create(&ctx, device);
    if (st != GS_OK) fail_exit(nullptr, st, "gs_create");
    // launch shapes from the autotuner's cache, when one is named
    if (const char *cache = std::getenv("GS_AUTOTUNE_CACHE")) {
        uint32_t n = 0;
        if (gs_autotune_load(ctx, cache, &n) != GS_OK)
            std::fprintf(stderr, "[warning] autotune cache %s: %s\n", cache, gs_last_error(ctx));
        else
            log_info("Launch shapes from " + std::string(cache) + ": " + std::to_string(n) + " entries.");
    }
    ctxs[device] = ctx;
    return ctx;
}

void fail_exit(gs_ctx *ctx, gs_status st, const char *what) {
    std::fprintf(stderr, "GPUassert: %s failed (status %d): %s\n", what, (int)st,
                 ctx ? gs_last_error(ctx) : "no device context");
    std::exit((int)st);
}

void run_model(int model, ParamsCarrier *pc) {
    log_info(std::string("Starting initialization of ") + model_name(model) + " simulation.");
    RunConfig cfg;
    std::string err;
    if (!config_from_carrier(pc, model, &cfg, &err)) {
        log_error("Invalid parameters: " + err);
        return;
    }
    std::string destination = pc->getString("destination", "");
    if (destination.empty()) {
        destination = directory_name(pc);
        log_info("Destination is not specified - using generated name for destination: " + destination);
    }
    if (!create_directory(model_directory(model), destination)) {
        log_error(std::string("Directory for ") + model_name(model) + " simulations cannot be created.");
        return;
    }
    gs_ctx *ctx = context(cfg.device);
    gs_spectrum spec;
    const gs_params &p = cfg.params;
    if (cfg.write_log) {
        // Every event fits: B-p and 2D models emit at most one per particle;
        // for F-p the call reports the total and runs again with room for it.
        uint64_t cap = model == GS_MODEL_1D_FP ? std::max<uint64_t>(p.n_particles / 8, 1 << 16)
                                               : std::max<uint64_t>(p.n_particles, 1);
        std::vector<gs_exit_record> recs(cap);
        uint64_t n = 0;
        gs_status st = gs_run_plain(ctx, &p, &spec, recs.data(), cap, &n);
        if (st != GS_OK) fail_exit(ctx, st, "gs_run_plain");
        if (n > cap) {
            recs.resize(n);
            st = gs_run_plain(ctx, &p, &spec, recs.data(), n, &n);
            if (st != GS_OK) fail_exit(ctx, st, "gs_run_plain");
        }
        recs.resize(n);
        log_info("Detected " + std::to_string(n) + " particles.");
        if (!write_log_dat("log.dat", model, recs, cfg.theta_inj_f)) log_error("Cannot write log.dat.");
    } else {
        const gs_status st = gs_run_plain(ctx, &p, &spec, nullptr, 0, nullptr);
        if (st != GS_OK) fail_exit(ctx, st, "gs_run_plain");
    }
    write_histogram("gs_histogram", p, spec, pc->getInt("csv", 0) != 0);
    write_simulation_report(pc, cfg);
    log_info("Simulation ended.");
}

void run_algorithm(int model, ParamsCarrier *pc) {
    if (pc->getInt("run_simulation", 1)) run_model(model, pc);
    if (model == GS_MODEL_1D_BP) results_1d_bp(pc);
    else if (model == GS_MODEL_1D_FP) results_1d_fp(pc);
    else results_2d(pc);
    if (pc->getInt("remove_log_files_after_simulation", 1)) unlink("log.dat");
}

}  // namespace gsdrop

void OneDimensionFpGpuModel::prepareAndRunSimulation(ParamsCarrier *singleTone) {
    gsdrop::run_model(GS_MODEL_1D_FP, singleTone);
}
void OneDimensionBpGpuModel::prepareAndRunSimulation(ParamsCarrier *singleTone) {
    gsdrop::run_model(GS_MODEL_1D_BP, singleTone);
}
void SolarPropLikeGpuModel::prepareAndRunSimulation(ParamsCarrier *singleTone) {
    gsdrop::run_model(GS_MODEL_2D_SOLARPROP, singleTone);
}
void GeliosphereGpuModel::prepareAndRunSimulation(ParamsCarrier *singleTone) {
    gsdrop::run_model(GS_MODEL_2D_GELIOSPHERE, singleTone);
}
