// gs_batch.cpp - batch CSV rows through one gs_run_batch call.
//
// The reference runs the rows of its batch file one after another, each as
// a full simulation (Algorithm/src/BatchRun.cpp:26-79).  Here every row is
// turned into a parameter block the way the reference's per-row parsing
// fills the carrier (BatchRun.cpp:282-342: dt as given, K0 in cm^2/s,
// V in km/s, N in millions, r and theta of injection, name as destination),
// all rows run together on the device, and each row's histogram and report
// are written to the directory its own run would have used.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>

#include "gs_dropin.hpp"

namespace gsdrop {

namespace {

std::vector<std::string> split_csv(const std::string &line) {
    std::vector<std::string> out;
    std::string cur;
    bool quoted = false;
    for (char ch : line) {
        if (ch == '"') quoted = !quoted;
        else if (ch == ',' && !quoted) { out.push_back(cur); cur.clear(); }
        else if (ch != '\r') cur += ch;
    }
    out.push_back(cur);
    return out;
}

double cell_num(const std::vector<std::string> &cells, int col) {
    if (col < 0 || col >= (int)cells.size() || cells[col].empty()) return -1.0;
    char *end = nullptr;
    const double v = std::strtod(cells[col].c_str(), &end);
    return (end && *end == 0) ? v : -1.0;   // unparsable cells read as -1, as the reference's converter does
}

}  // namespace

bool read_batch_csv(const std::string &path, std::vector<BatchRow> *rows, std::string *err) {
    std::ifstream in(path);
    if (!in.good()) {
        *err = "Cannot open input batch file.";
        return false;
    }
    std::string line;
    if (!std::getline(in, line)) {
        *err = "Empty batch file.";
        return false;
    }
    const std::vector<std::string> head = split_csv(line);
    auto col = [&](const char *name) {
        for (size_t i = 0; i < head.size(); ++i)
            if (head[i] == name) return (int)i;
        return -1;
    };
    const int cy = col("year"), cm = col("month"), cK = col("K0"), cV = col("V"), cdt = col("dt"),
              cN = col("N"), cr = col("r"), ct = col("theta"), cs = col("pathToCustomSettingsFile"),
              cn = col("name"), cmod = col("model");
    if (cmod < 0) {
        *err = "Column model is not defined.";
        return false;
    }
    while (std::getline(in, line)) {
        if (line.empty() || line == "\r") continue;
        const std::vector<std::string> c = split_csv(line);
        BatchRow r;
        r.year = (int)cell_num(c, cy);
        r.month = (int)cell_num(c, cm);
        r.K0 = cell_num(c, cK);
        r.V = cell_num(c, cV);
        r.dt = cell_num(c, cdt);
        r.N = (long)cell_num(c, cN);
        r.r = cell_num(c, cr);
        r.theta = cell_num(c, ct);
        if (cs >= 0 && cs < (int)c.size()) r.settings = c[cs];
        if (cn >= 0 && cn < (int)c.size()) r.name = c[cn];
        if (cmod < (int)c.size()) r.model = c[cmod];
        rows->push_back(r);
    }
    return true;
}

int run_batch_rows(ParamsCarrier *pc, const std::vector<BatchRow> &rows) {
    const std::set<std::string> keep = {"default_settings_file", "currentApplicationPath", "csv", "isCpu",
                                        "inputBatchFile", "gs_seed", "gs_device", "gs_particles",
                                        "gs_solarprop_gpu_compat"};
    std::vector<gs_params> sims;
    std::vector<RunConfig> cfgs;
    std::vector<std::string> dests;
    std::vector<int> models;
    const std::string seed0 = pc->getString("gs_seed", "");
    for (size_t i = 0; i < rows.size(); ++i) {
        const BatchRow &b = rows[i];
        pc->eraseAllItems(keep);
        const int model = model_from_name(b.model);
        if (model < 0) {
            std::fprintf(stderr, "[error] Selected model for simulation %zu is not supported in batch mode.\n", i + 1);
            continue;
        }
        // Checks of BatchRun::parseValues / InputValidation.
        if (b.dt != -1.0 && (b.dt < 3.0 || b.dt > 5000.0)) {
            std::fprintf(stderr, "[error] dt is out of range!(3-5000); skipping simulation %zu.\n", i + 1);
            continue;
        }
        if (b.K0 != -1.0 && b.K0 < 0.0) {
            std::fprintf(stderr, "[error] K0 is out of range!(>0); skipping simulation %zu.\n", i + 1);
            continue;
        }
        if (b.N != -1 && b.N <= 0) {
            std::fprintf(stderr, "[error] Number of test particles must be greater than 0; skipping %zu.\n", i + 1);
            continue;
        }
        if (b.V != -1.0 && (b.V < 100.0 || b.V > 1500.0)) {
            std::fprintf(stderr, "[error] V is out of range!(100-1500 km/s); skipping simulation %zu.\n", i + 1);
            continue;
        }
        // BatchRun.cpp:262-273
        if ((b.month != -1) != (b.year != -1)) {
            std::fprintf(stderr, "[error] Both month and year must be set; skipping simulation %zu.\n", i + 1);
            continue;
        }
        const bool dated = b.month != -1 && b.year != -1;
        if (dated && (b.K0 != -1.0 || b.V != -1.0)) {
            std::fprintf(stderr, "[error] Simulation dates cannot be combined with K0 or V; skipping simulation %zu.\n",
                         i + 1);
            continue;
        }
        pc->putString("model", b.model);
        if (b.dt != -1.0) pc->putFloat("dt", (float)b.dt);
        if (b.K0 != -1.0) {
            pc->putFloat("K0", (float)((float)b.K0 * 4.4683705e-27));
            pc->putInt("K0_entered_by_user", 1);
        }
        if (b.N != -1) pc->putInt("millions", (int)b.N);
        if (b.V != -1.0) pc->putFloat("V", (float)((float)b.V * 6.68458712e-9));
        // The row's settings file, or the batch defaults, with its r and theta
        // (BatchRun.cpp:195-216,335-336).
        Settings st = batch_default_settings();
        if (!b.settings.empty() && access(b.settings.c_str(), R_OK) == 0) {
            std::string serr;
            if (!read_settings(b.settings, &st, &serr)) {
                std::fprintf(stderr, "[error] %s; skipping simulation %zu.\n", serr.c_str(), i + 1);
                continue;
            }
        }
        if (b.r > 0.0) st.r_injection = b.r;
        if (b.theta > 0.0) st.theta_injection = b.theta;
        put_settings(pc, st);
        if (!b.name.empty() && b.name.find_first_not_of(' ') != std::string::npos) pc->putString("destination", b.name);
        if (dated) {   // BatchRun.cpp:337-340, after the row's other values
            std::string terr;
            if (!month_year_check(pc, model, b.year, b.month, pc->getString("currentApplicationPath", "./"), &terr))
                std::fprintf(stderr, "[error] %s (simulation %zu)\n", terr.c_str(), i + 1);
        }
        // Distinct particle streams per row unless the caller fixed a seed.
        if (!seed0.empty()) pc->putString("gs_seed", std::to_string(std::strtoull(seed0.c_str(), nullptr, 10) + i));
        RunConfig cfg;
        std::string err;
        if (!config_from_carrier(pc, model, &cfg, &err)) {
            std::fprintf(stderr, "[error] Unable to parse values for %zu. simulation: %s\n", i + 1, err.c_str());
            continue;
        }
        std::string dest = pc->getString("destination", "");
        if (dest.empty()) dest = directory_name(pc) + "_" + std::to_string(i + 1);
        sims.push_back(cfg.params);
        cfgs.push_back(cfg);
        dests.push_back(dest);
        models.push_back(model);
        if (!seed0.empty()) pc->putString("gs_seed", seed0);
    }
    if (sims.empty()) return 0;
    gs_ctx *ctx = context(cfgs[0].device);
    std::vector<gs_spectrum> outs(sims.size());
    const gs_status st = gs_run_batch(ctx, sims.data(), (uint32_t)sims.size(), outs.data());
    if (st != GS_OK) fail_exit(ctx, st, "gs_run_batch");
    for (size_t i = 0; i < sims.size(); ++i) {
        if (!create_directory(model_directory(models[i]), dests[i])) {
            std::fprintf(stderr, "[error] Directory for simulation %zu cannot be created.\n", i + 1);
            continue;
        }
        write_histogram("gs_histogram", sims[i], outs[i], pc->getInt("csv", 0) != 0);
        FILE *f = std::fopen("simulation_log.txt", "w");
        if (f) {
            const gs_params &p = sims[i];
            std::fprintf(f, "Model: %s\n---------------------\n", model_name(models[i]));
            std::fprintf(f, "dt: %g\nK0: %g\nV: %g\n", p.dt, p.K0, p.V);
            std::fprintf(f, "r_injection: %g\ntheta_injection: %g\n", p.r_inj, p.theta_inj);
            std::fprintf(f, "Engine: gfx950 SDE trajectory engine (FP64, Philox4x32-10), batched\n");
            std::fprintf(f, "Seed: %llu\nTest particles: %llu\n", (unsigned long long)p.seed,
                         (unsigned long long)p.n_particles);
            std::fclose(f);
        }
        if (chdir("../../..") == -1) {
            std::fprintf(stderr, "[error] Could not change directory\n");
            break;
        }
    }
    return (int)sims.size();
}

}  // namespace gsdrop
