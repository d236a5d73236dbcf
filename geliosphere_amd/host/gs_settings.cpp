// gs_settings.cpp - the settings file of a run: the reference's TOML schema
// (Input/src/TomlSettings.cpp:9-52; Settings.toml) read by a parser for the
// subset it uses - [table] headers, key = value with floats, integers and
// booleans, # comments - and the table a batch row starts from when it names
// no readable file (Algorithm/src/BatchRun.cpp:195-252).
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>

#include "gs_dropin.hpp"

namespace gsdrop {

namespace {

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// "table.key" -> raw value text
bool parse_toml(const std::string &path, std::map<std::string, std::string> *kv, std::string *err) {
    std::ifstream in(path);
    if (!in) {
        *err = "cannot open " + path;
        return false;
    }
    std::string line, table;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        bool quoted = false;
        for (size_t i = 0; i < line.size(); ++i) {   // strip a comment outside quotes
            if (line[i] == '"') quoted = !quoted;
            if (line[i] == '#' && !quoted) { line.resize(i); break; }
        }
        line = trim(line);
        if (line.empty()) continue;
        if (line.front() == '[') {
            if (line.back() != ']') {
                *err = path + ":" + std::to_string(lineno) + ": bad table header";
                return false;
            }
            table = trim(line.substr(1, line.size() - 2));
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) {
            *err = path + ":" + std::to_string(lineno) + ": expected key = value";
            return false;
        }
        (*kv)[table + "." + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return true;
}

bool number(const std::map<std::string, std::string> &kv, const std::string &key, double *v, std::string *err) {
    const auto it = kv.find(key);
    if (it == kv.end()) {
        *err = "settings: missing " + key;
        return false;
    }
    std::string t;
    for (char c : it->second)
        if (c != '_') t += c;   // TOML digit separators
    char *end = nullptr;
    *v = std::strtod(t.c_str(), &end);
    if (t.empty() || *end != '\0') {
        *err = "settings: " + key + " is not a number";
        return false;
    }
    return true;
}

bool boolean(const std::map<std::string, std::string> &kv, const std::string &key, bool *v, std::string *err) {
    const auto it = kv.find(key);
    if (it == kv.end()) {
        *err = "settings: missing " + key;
        return false;
    }
    if (it->second == "true") *v = true;
    else if (it->second == "false") *v = false;
    else {
        *err = "settings: " + key + " is not a boolean";
        return false;
    }
    return true;
}

}  // namespace

Settings batch_default_settings() {
    Settings s;
    s.use_uniform_injection = true;   // as BatchRun::getDefaultSettingsFile writes it
    return s;
}

bool read_settings(const std::string &path, Settings *s, std::string *err) {
    std::map<std::string, std::string> kv;
    if (!parse_toml(path, &kv, err)) return false;
    return number(kv, "default_values.K0", &s->K0, err) && number(kv, "default_values.V", &s->V, err) &&
           number(kv, "default_values.dt", &s->dt, err) &&
           number(kv, "default_values.r_injection", &s->r_injection, err) &&
           number(kv, "default_values.uniform_energy_injection_maximum", &s->injection_max, err) &&
           number(kv, "2d_models_common_settings.theta_injection", &s->theta_injection, err) &&
           boolean(kv, "2d_models_common_settings.use_uniform_injection", &s->use_uniform_injection, err) &&
           number(kv, "SolarProp_like_model_settings.SolarProp_ratio", &s->solarprop_ratio, err) &&
           number(kv, "Geliosphere_model_settings.Geliosphere_ratio", &s->geliosphere_ratio, err) &&
           number(kv, "Geliosphere_model_settings.K0_ratio", &s->K0_ratio, err) &&
           number(kv, "Geliosphere_model_settings.C_delta", &s->C_delta, err) &&
           number(kv, "Geliosphere_model_settings.default_tilt_angle", &s->default_tilt_angle, err) &&
           boolean(kv, "advanced_settings.remove_log_files_after_simulation", &s->remove_log_files, err);
}

void put_settings(ParamsCarrier *pc, const Settings &s) {
    // TomlSettings::parseFromSettings: doubles stored as floats
    pc->putFloat("K0_default", (float)(s.K0 * 4.4683705e-27));
    pc->putFloat("V_default", (float)(s.V * 6.68458712e-9));
    pc->putFloat("dt_default", (float)s.dt);
    pc->putFloat("r_injection", (float)s.r_injection);
    pc->putFloat("uniform_energy_injection_maximum", (float)s.injection_max);
    pc->putFloat("theta_injection", (float)s.theta_injection);
    pc->putInt("use_uniform_injection", s.use_uniform_injection ? 1 : 0);
    pc->putFloat("solarPropRatio", (float)s.solarprop_ratio);
    pc->putFloat("geliosphereRatio", (float)s.geliosphere_ratio);
    pc->putFloat("K0_ratio", (float)s.K0_ratio);
    pc->putFloat("C_delta", (float)s.C_delta);
    pc->putFloat("default_tilt_angle", (float)s.default_tilt_angle);
    pc->putInt("remove_log_files_after_simulation", s.remove_log_files ? 1 : 0);
}

}  // namespace gsdrop
