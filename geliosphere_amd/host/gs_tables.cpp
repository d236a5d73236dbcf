// gs_tables.cpp - measured solar-cycle values for a month and year: the
// diffusion coefficient K0 with the solar wind speed (1D models) or the
// current-sheet tilt and field polarity (2D models), looked up in the
// reference's CSV tables in the application directory
// (Input/src/MeasureValuesTransformation.cpp:14-115,
// Input/src/InputValidation.cpp:22-37,130-151).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gs_dropin.hpp"

namespace gsdrop {

const char *table_name(int model) {
    if (model == GS_MODEL_2D_SOLARPROP) return "SolarProp_K0_phi_table.csv";
    if (model == GS_MODEL_2D_GELIOSPHERE) return "Geliosphere_K0_phi_table.csv";
    return "K0_phi_table.csv";
}

long carrington_rotation(int month, int year) {
    // Rotation 1642 starts on 1976-05-27; the 15th of the month counts, in
    // rotations of 2356560 s, rounded up from a fraction above 0.5499.
    std::tm start{};
    start.tm_year = 1976 - 1900;
    start.tm_mon = 5 - 1;
    start.tm_mday = 27;
    start.tm_isdst = -1;
    std::tm end{};
    end.tm_year = year - 1900;
    end.tm_mon = month - 1;
    end.tm_mday = 15;
    end.tm_isdst = -1;
    const double seconds = std::difftime(std::mktime(&end), std::mktime(&start));
    double whole = 0.0;
    double rot = seconds / 2356560.0;
    const double frac = std::modf(rot, &whole);
    rot = frac <= 0.5499 ? std::trunc(rot) : std::ceil(rot);
    return (long)rot + 1642;
}

std::string table_row_id(int model, int month, int year) {
    if (model == GS_MODEL_2D_SOLARPROP || model == GS_MODEL_2D_GELIOSPHERE)
        return std::to_string(carrington_rotation(month, year));
    // 1D tables: "<year>.<fraction of the year at the month's middle>"
    static const char *const kMonth[12] = {"04", "12", "20", "29", "37", "45", "54", "62", "70", "79", "87", "95"};
    return std::to_string(year) + "." + (month >= 1 && month <= 12 ? kMonth[month - 1] : "NaN");
}

namespace {

std::vector<std::string> split_csv(const std::string &line) {
    std::vector<std::string> out;
    std::string cell;
    std::istringstream ss(line);
    while (std::getline(ss, cell, ',')) {
        if (!cell.empty() && cell.back() == '\r') cell.pop_back();
        out.push_back(cell);
    }
    if (!line.empty() && line.back() == ',') out.push_back("");
    return out;
}

}  // namespace

bool table_lookup(const std::string &path, const std::string &row_id, const std::vector<std::string> &columns,
                  std::vector<std::string> *values, std::string *err) {
    std::ifstream in(path);
    if (!in) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::string line;
    if (!std::getline(in, line)) {
        if (err) *err = path + " is empty";
        return false;
    }
    const std::vector<std::string> head = split_csv(line);
    std::vector<size_t> idx;
    for (const std::string &c : columns) {
        size_t k = 0;
        while (k < head.size() && head[k] != c) ++k;
        if (k == head.size()) {
            if (err) *err = "column " + c + " missing in " + path;
            return false;
        }
        idx.push_back(k);
    }
    while (std::getline(in, line)) {
        const std::vector<std::string> cells = split_csv(line);
        if (cells.empty() || cells[0] != row_id) continue;
        values->clear();
        for (size_t k : idx) values->push_back(k < cells.size() ? cells[k] : "");
        return true;
    }
    if (err) *err = "Combination for entered date was not found in input table.";
    return false;
}

bool month_year_check(ParamsCarrier *pc, int model, int year, int month, const std::string &dir, std::string *err) {
    pc->putInt("month_option", month);
    pc->putInt("year_option", year);
    const std::string path = dir + table_name(model);
    const std::string row = table_row_id(model, month, year);
    const bool two_d = model == GS_MODEL_2D_SOLARPROP || model == GS_MODEL_2D_GELIOSPHERE;
    std::vector<std::string> v;
    if (!table_lookup(path, row, two_d ? std::vector<std::string>{"k0_au2/s", "tilt_angle", "polarity"}
                                        : std::vector<std::string>{"k0_au2/s", "V"},
                      &v, err))
        return false;
    // The reference reads the cells as float (and int for the polarity) and
    // converts V from km/s to AU/s in double before storing a float.
    pc->putFloat("K0", std::strtof(v[0].c_str(), nullptr));
    if (two_d) {
        pc->putFloat("tilt_angle", std::strtof(v[1].c_str(), nullptr));
        pc->putInt("polarity", (int)std::strtol(v[2].c_str(), nullptr, 10));
    } else {
        pc->putFloat("V", (float)(std::strtof(v[1].c_str(), nullptr) * 6.68458712e-9));
    }
    return true;
}

}  // namespace gsdrop
