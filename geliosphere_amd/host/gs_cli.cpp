// gs_cli.cpp - command-line driver for the drop-in GPU models.
//
// Accepts the reference CLI's model and parameter flags
// (Input/src/ParseParams.cpp:29-62) and converts them into carrier keys the
// way the reference's InputValidation does (K0 in cm^2/s * 4.4683705e-27,
// V in km/s * 6.68458712e-9, N in millions), then runs the model's
// Algorithm step: simulation, Results pass, removal of log.dat.
//
//   gs_cli (-F|-B|-E|-T|--custom-model NAME|-b FILE) [-N millions] [-K K0]
//          [-V V] [-d dt] [-m month -y year] [-s settings.toml]
//          [-p destination] [-c] [--evaluation log.dat] [--cpu-only]
//          [--r AU] [--theta deg] [--tilt deg] [--polarity +-1]
//          [--seed N] [--particles N] [--keep-log] [--no-log]
//          [--solarprop-gpu-compat] [--device N] [--print-params]
//
// Settings come from -s FILE (the reference's TOML schema); without -s the
// reference reads no settings file (its default-path check tests the empty
// option string, ParseParams.cpp:126-133), so neither does gs_cli.  -m/-y take
// K0 and V (1D) or K0, tilt and polarity (2D) from the measured-value table
// next to the executable; they exclude -K and -V, and -m needs -y.
// --print-params prints the parameter block the model would run with (JSON)
// and exits without touching a device.
//
// The model flags -F, -B, -E, -T, --custom-model and -b exclude each other
// (ParseParams.cpp:54-59).  --custom-model NAME runs the model the
// reference's factory returns for NAME (Factory/src/CosmicFactory.cpp:8-31;
// main.cpp:34-39): one of the four model names, or "batch run" with the
// default batch file; any other name is an error.  --cpu-only selects the
// reference's CPU models (ParseParams.cpp:36,136-139), which are not part of
// this GPU drop-in: it is rejected with a message.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gs_dropin.hpp"

namespace {

void usage() {
    std::fprintf(stderr,
                 "usage: gs_cli (-F|-B|-E|-T|--custom-model NAME|-b FILE) [-N millions] [-K K0_cm2s]\n"
                 "              [-V V_kms] [-d dt] [-m month -y year] [-s settings.toml] [-p destination]\n"
                 "              [-c] [--evaluation LOG] [--cpu-only]\n"
                 "              [--r AU] [--theta deg]\n"
                 "              [--tilt deg] [--polarity 1|-1] [--seed N] [--particles N]\n"
                 "              [--keep-log] [--no-log] [--solarprop-gpu-compat] [--device N]\n"
                 "              [--print-params]\n");
}

// The fields of a parameter block as JSON members, closing the object.
void print_params_json(const gs_params &p) {
    std::printf("\"model\": %d, \"injection\": %d, \"flags\": %u, \"K0\": %.17g, \"V\": %.17g, "
                "\"dt\": %.17g, \"r_inj\": %.17g, \"theta_inj\": %.17g, \"tilt\": %.17g, "
                "\"polarity\": %.17g, \"ratio\": %.17g, \"delta0\": %.17g, \"T_min\": %.17g, "
                "\"T_max\": %.17g, \"grid_n\": %u, \"n_particles\": %llu, \"seed\": %llu, "
                "\"n_bins\": %u, \"n_weights\": %u, \"weights\": [%d, %d, %d, %d]}\n",
                p.model, p.injection, p.flags, p.K0, p.V, p.dt, p.r_inj, p.theta_inj, p.tilt, p.polarity, p.ratio,
                p.delta0, p.T_min, p.T_max, p.grid_n, (unsigned long long)p.n_particles, (unsigned long long)p.seed,
                p.n_bins, p.n_weights, p.weights[0], p.weights[1], p.weights[2], p.weights[3]);
}

}  // namespace

int main(int argc, char **argv) {
    ParamsCarrier *pc = ParamsCarrier::instance();
    int model = -1;
    std::string batch;
    std::string custom_model;
    bool print_params = false;
    bool cpu_only = false;
    std::vector<std::string> model_flags;   // -F -B -E -T --custom-model -b, as given
    int month = -1, year = -1;
    std::string settings;
    bool settings_set = false;
    // The carrier keeps the first value put under a key (ParamsCarrier's map
    // insert), so flags given here win over the settings file put below.
    bool k_set = false, v_set = false, other_set = false;
    // the application directory: argv[0] up to its last '/' (ParseParams.cpp getApplicationPath)
    const std::string a0 = argv[0];
    const std::string app_dir = a0.rfind('/') == std::string::npos ? "" : a0.substr(0, a0.rfind('/') + 1);
    pc->putString("currentApplicationPath", app_dir);
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-F" || a == "--forward") { model = GS_MODEL_1D_FP; model_flags.push_back("--forward"); }
        else if (a == "-B" || a == "--backward") { model = GS_MODEL_1D_BP; model_flags.push_back("--backward"); }
        else if (a == "-E" || a == "--solarprop-like-model") {
            model = GS_MODEL_2D_SOLARPROP;
            model_flags.push_back("--solarprop-like-model");
        } else if (a == "-T" || a == "--geliosphere-2d-model") {
            model = GS_MODEL_2D_GELIOSPHERE;
            model_flags.push_back("--geliosphere-2d-model");
        } else if (a == "--custom-model") { custom_model = val(); model_flags.push_back("--custom-model"); }
        else if (a == "-b" || a == "--batchrun") { batch = val(); model_flags.push_back("--batchrun"); }
        else if (a == "--cpu-only") cpu_only = true;
        else if (a == "-c" || a == "--csv") pc->putInt("csv", 1);
        else if (a == "-N" || a == "--number-of-test-particles") {
            const int n = std::atoi(val());
            if (n <= 0) { std::fprintf(stderr, "Number of test particles must be greater than 0!\n"); return 2; }
            pc->putInt("millions", n);
            other_set = true;
        } else if (a == "-s" || a == "--settings") {
            settings = val();
            settings_set = true;
        } else if (a == "-m" || a == "--month") {
            month = std::atoi(val());
        } else if (a == "-y" || a == "--year") {
            year = std::atoi(val());
        } else if (a == "-K" || a == "--K0") {
            const float K = (float)std::atof(val());
            if (K < 0.0f) { std::fprintf(stderr, "K0 is out of range!(>0)\n"); return 2; }
            pc->putFloat("K0", (float)(K * 4.4683705e-27));
            pc->putInt("K0_entered_by_user", 1);
            k_set = true;
        } else if (a == "-V" || a == "--V") {
            const float V = (float)std::atof(val());
            if (V < 100.0f || V > 1500.0f) { std::fprintf(stderr, "V is out of range!(100-1500 km/s)\n"); return 2; }
            pc->putFloat("V", (float)(V * 6.68458712e-9));
            v_set = true;
        } else if (a == "-d" || a == "--dt") {
            const float dt = (float)std::atof(val());
            if (dt < 3.0f || dt > 5000.0f) { std::fprintf(stderr, "dt is out of range!(3-5000)\n"); return 2; }
            pc->putFloat("dt", dt);
            other_set = true;
        } else if (a == "-p" || a == "--path") pc->putString("destination", val());
        else if (a == "--evaluation") {
            pc->putInt("run_simulation", 0);
            pc->putString("pathToLogFile", val());
        } else if (a == "--r") pc->putFloat("r_injection", (float)std::atof(val()));
        else if (a == "--theta") pc->putFloat("theta_injection", (float)std::atof(val()));
        else if (a == "--tilt") pc->putFloat("tilt_angle", (float)std::atof(val()));
        else if (a == "--polarity") pc->putInt("polarity", std::atoi(val()));
        else if (a == "--seed") pc->putString("gs_seed", val());
        else if (a == "--particles") pc->putInt("gs_particles", std::atoi(val()));
        else if (a == "--keep-log") pc->putInt("remove_log_files_after_simulation", 0);
        else if (a == "--no-log") pc->putInt("gs_write_log", 0);
        else if (a == "--solarprop-gpu-compat") pc->putInt("gs_solarprop_gpu_compat", 1);
        else if (a == "--device") pc->putInt("gs_device", std::atoi(val()));
        else if (a == "--print-params") print_params = true;
        else {
            usage();
            return 2;
        }
    }
    // option relations of the reference CLI (ParseParams.cpp:51-62)
    if (model_flags.size() > 1) {
        std::fprintf(stderr, "%s excludes %s\n", model_flags[1].c_str(), model_flags[0].c_str());
        return 2;
    }
    if (cpu_only) {
        std::fprintf(stderr, "[error] --cpu-only selects the reference's CPU models (CpuImplementations/), which "
                             "this GPU drop-in does not provide; build the reference with GPU_ENABLED=0 for them.\n");
        return 2;
    }
    if (!custom_model.empty()) {
        // the reference's factory: the four model names and "batch run"
        if (custom_model == "batch run") batch = "../geliosphere_batch_input_paramaters.csv";
        else model = gsdrop::model_from_name(custom_model);
    }
    const bool date = month != -1 || year != -1;
    if ((k_set || v_set) && date) { std::fprintf(stderr, "-K/-V exclude -m/-y\n"); return 2; }
    if (month != -1 && year == -1) { std::fprintf(stderr, "-m requires -y\n"); return 2; }
    if (settings_set && access(settings.c_str(), F_OK) == 0) {
        gsdrop::Settings st;
        std::string err;
        if (!gsdrop::read_settings(settings, &st, &err)) {
            std::fprintf(stderr, "[error] %s\n", err.c_str());
            return 1;
        }
        gsdrop::put_settings(pc, st);
    } else {
        std::fprintf(stderr, settings_set ? "[warning] No settings file exists on entered path.\n"
                                          : "[warning] No settings file exists on default path.\n");
    }
    if (!custom_model.empty() && batch.empty() && model < 0) {
        std::fprintf(stderr, "[error] Selected custom model is not defined in factory.\n");
        return 255;
    }
    if (!batch.empty()) {
        if (k_set || v_set || other_set || date) {
            std::fprintf(stderr, "-b excludes -d, -N, -K, -V, -m and -y\n");
            return 2;
        }
        std::vector<gsdrop::BatchRow> rows;
        std::string err;
        if (!gsdrop::read_batch_csv(batch, &rows, &err)) {
            std::fprintf(stderr, "[error] %s\n", err.c_str());
            return 1;
        }
        if (print_params) {
            // the parameter block and directory of every row that would run
            const gsdrop::PreparedBatch B = gsdrop::prepare_batch_rows(pc, rows);
            for (size_t i = 0; i < B.cfg.size(); ++i) {
                std::printf("{\"row\": %zu, \"destination\": \"%s\", ", B.row[i] + 1, B.dest[i].c_str());
                print_params_json(B.cfg[i].params);
            }
            return 0;
        }
        const int n = gsdrop::run_batch_rows(pc, rows);
        std::fprintf(stderr, "[info] %d of %zu batch simulations ran.\n", n, rows.size());
        return n == (int)rows.size() ? 0 : 1;
    }
    if (model < 0) {
        std::fprintf(stderr, "At least one model must be selected!\n");
        usage();
        return 2;
    }
    pc->putString("model", gsdrop::model_name(model));
    if (month != -1 && year != -1) {
        std::string err;
        if (!gsdrop::month_year_check(pc, model, year, month, app_dir, &err))
            std::fprintf(stderr, "[error] %s\n", err.c_str());   // the run goes on without the table values
    }
    if (print_params) {
        gsdrop::RunConfig cfg;
        std::string err;
        if (!gsdrop::config_from_carrier(pc, model, &cfg, &err)) {
            std::fprintf(stderr, "[error] %s\n", err.c_str());
            return 1;
        }
        std::printf("{");
        print_params_json(cfg.params);
        return 0;
    }
    gsdrop::run_algorithm(model, pc);
    return 0;
}
