// gs_dropin.hpp - host side of the drop-in GPU models.
//
// The four adapter classes (OneDimensionFpGpuModel, OneDimensionBpGpuModel,
// SolarPropLikeGpuModel, GeliosphereGpuModel) map the reference's
// ParamsCarrier keys onto a gs_params block, run the simulation through the
// C ABI (include/gs_api.h) and leave the same side effects as the
// reference's GPU models: the working directory is
// results/<MODEL>/<destination>/, which holds log.dat (one line per exit or
// 1 AU crossing, in the reference's per-model format) and
// simulation_log.txt.  They also write gs_histogram.dat (.csv): the
// device-built 30-bin log spectrum with exact weighted sums.
//
// Everything else lives in namespace gsdrop so that the adapters can be
// linked into the reference application next to its own Utils/ and
// Algorithm/ code without symbol clashes.
#ifndef GS_DROPIN_HPP
#define GS_DROPIN_HPP

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gs_api.h"
#include "ParamsCarrier.hpp"

namespace gsdrop {

// Reference model names (CosmicFactory.cpp:8-31) <-> gs_model; -1 if unknown.
int model_from_name(const std::string &name);
const char *model_name(int model);
// results/<dir>/ per model (OneDimensionFpModel.cu:163, OneDimensionBpModel.cu:162,
// SolarPropLikeModel.cu:269, GeliosphereModel.cu:412).
const char *model_directory(int model);

// Options the reference does not have; all read from the carrier.
//   gs_seed (string, decimal u64)     Philox key; default: derived from the clock,
//                                     as the reference seeds from clock()
//   gs_particles (int)                exact particle count (default: "millions"
//                                     rounded up to whole 2^20 launches, as the
//                                     reference GPU path does)
//   gs_write_log (int, default 1)     write log.dat
//   gs_solarprop_gpu_compat (int, 0)  2D SolarProp-like as the reference GPU
//                                     kernel runs it: tilt 5.75 deg, theta
//                                     drift with polarity +1
//   gs_device (int, 0)                HIP device
struct RunConfig {
    gs_params params;
    uint64_t seed = 0;
    bool write_log = true;
    float theta_inj_f = 0.f;   // theta_injection as the reference prints it (float rad)
    int device = 0;
};

// gs_params from the carrier, following the reference's constant setup
// (CUDAKernel/src/CosmicConstants.cu:43-111, CpuImplementations/include/
// Constants.hpp:157-223).  Returns false with *err on invalid input.
bool config_from_carrier(ParamsCarrier *pc, int model, RunConfig *cfg, std::string *err);

// Working-directory contract (reference Utils/src/FileUtils.cpp:14-97).
bool create_directory(const std::string &model_dir, const std::string &destination);
std::string directory_name(ParamsCarrier *pc);
// simulation_log.txt (FileUtils.cpp:99-125) plus the engine and seed: its
// text, and the file in the working directory.
std::string simulation_report(ParamsCarrier *pc, const RunConfig &cfg, const char *engine);
void write_simulation_report(ParamsCarrier *pc, const RunConfig &cfg);

// log.dat lines in the reference's per-model formats
// (OneDimensionFpModel.cu:195-196, OneDimensionBpModel.cu:192-193,
// SolarPropLikeModel.cu:299, GeliosphereModel.cu:442).
bool write_log_dat(const char *path, int model, const std::vector<gs_exit_record> &recs,
                   float theta_inj_f);
// The device histogram: per bin edges, count, and per weight sum w, sum w^2,
// mean and its standard error.
bool write_histogram(const std::string &base, const gs_params &p, const gs_spectrum &s, bool csv);

// One context per process and device, created on first use.
gs_ctx *context(int device);
// Reference gpuErrchk convention (CudaErrorCheck.cuh:31-41): print and exit.
[[noreturn]] void fail_exit(gs_ctx *ctx, gs_status st, const char *what);

// The body shared by the four adapters' prepareAndRunSimulation.
void run_model(int model, ParamsCarrier *pc);

// Restated Results passes: read log.dat ("pathToLogFile") and write the
// reference's spectrum files in the working directory
// (Algorithm/src/OneDimensionBpResults.cpp:14-113, OneDimensionFpResults.cpp:
// 13-114, TwoDimensionBpResults.cpp:13-134; writer Utils/src/ResultsUtils.cpp).
bool results_1d_bp(ParamsCarrier *pc);
bool results_1d_fp(ParamsCarrier *pc);
bool results_2d(ParamsCarrier *pc);
// The Algorithm step for `model`: run (unless run_simulation == 0), Results,
// then remove log.dat unless remove_log_files_after_simulation == 0
// (Algorithm/src/OneDimensionBpAlgorithm.cpp:7-35).
void run_algorithm(int model, ParamsCarrier *pc);

// One row of the reference batch CSV (Algorithm/include/BatchRun.hpp:29-42).
struct BatchRow {
    int year = -1, month = -1;
    double K0 = -1, V = -1, dt = -1, r = -1, theta = -1;
    long N = -1;
    std::string settings, name, model;
};
// Rows of a batch CSV with the header year,month,K0,V,dt,N,r,theta,
// pathToCustomSettingsFile,name,model (README.md:39); empty cells are -1.
bool read_batch_csv(const std::string &path, std::vector<BatchRow> *rows, std::string *err);
// Rows turned into runs as the reference's per-row parsing fills the carrier
// (BatchRun.cpp:254-342); rows that fail a check are skipped with a message.
struct PreparedBatch {
    std::vector<size_t> row;           // index of the row in the file
    std::vector<int> model;
    std::vector<RunConfig> cfg;
    std::vector<std::string> dest;     // results/<MODEL>/<dest>/
    std::vector<std::string> report;   // simulation_log.txt
    std::vector<bool> remove_log;      // remove_log_files_after_simulation
};
PreparedBatch prepare_batch_rows(ParamsCarrier *pc, const std::vector<BatchRow> &rows);
// All rows in one gs_run_batch_records call; each row then leaves in
// results/<MODEL>/<dest>/ what its own run would (BatchRun.cpp:50-78 runs
// each row's Algorithm): log.dat (removed after the Results pass unless
// remove_log_files_after_simulation = 0), the Results pass's spectrum files,
// simulation_log.txt, and gs_histogram.dat.  Returns the number of rows run.
int run_batch_rows(ParamsCarrier *pc, const std::vector<BatchRow> &rows);

// The settings file's values (Settings.toml schema, Input/src/TomlSettings.cpp:
// 9-52); defaults are the shipped file's.
struct Settings {
    double K0 = 5e22, V = 400.0, dt = 50.0, r_injection = 1.0, injection_max = 101.0;
    double theta_injection = 90.0;
    bool use_uniform_injection = false;
    double solarprop_ratio = 0.02, geliosphere_ratio = 0.2, K0_ratio = 1.0, C_delta = 8.7e-5;
    double default_tilt_angle = 0.1;
    bool remove_log_files = true;
};
// Reads a settings file (the TOML subset the schema uses); false, with the
// reason in *err, when it cannot be read or a key is missing.
bool read_settings(const std::string &path, Settings *s, std::string *err);
// What a batch row uses when it names no readable settings file
// (Algorithm/src/BatchRun.cpp:218-252: the shipped values, but
// use_uniform_injection = true).
Settings batch_default_settings();
// Puts the values into the carrier as TomlSettings::parseFromSettings does.
void put_settings(ParamsCarrier *pc, const Settings &s);

// Measured values for a month and year (Input/src/MeasureValuesTransformation.cpp,
// InputValidation.cpp:22-37,130-151): the table file of `model`, the row key
// (1D: "<year>.<month code>", 2D: the Carrington rotation of the 15th of the
// month), and the lookup that puts K0 and V (1D) or K0, tilt_angle and
// polarity (2D) into the carrier, with month_option / year_option.  `dir` is
// the application directory (ending in '/').  False, with the reason in
// *err, when the file, a column or the row is missing.
const char *table_name(int model);
long carrington_rotation(int month, int year);
std::string table_row_id(int model, int month, int year);
bool table_lookup(const std::string &path, const std::string &row_id, const std::vector<std::string> &columns,
                  std::vector<std::string> *values, std::string *err);
bool month_year_check(ParamsCarrier *pc, int model, int year, int month, const std::string &dir, std::string *err);

}  // namespace gsdrop

#endif
