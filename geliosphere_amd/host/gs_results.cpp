// gs_results.cpp - the Results passes: log.dat -> spectrum files.
//
// Restates Algorithm/src/OneDimensionBpResults.cpp:14-113,
// OneDimensionFpResults.cpp:13-114 and TwoDimensionBpResults.cpp:13-134 with
// the file formats of Utils/src/ResultsUtils.cpp:16-126, so that the files
// a standalone build writes are the ones the reference writes from the same
// log.dat.  Behaviour kept on purpose because downstream tools read these
// files as the reference writes them:
//   * the last line of log.dat is not read (the passes read
//     countLines(log.dat) - 1 records);
//   * the "logbin" file prints the first 24 entries of the log spectrum
//     divided by the widths of a different log grid (10^(k/5 - 3));
//   * the 1e2/1e3/4e2 files print three columns although the CSV header
//     names four.
// One reference defect is not reproduced: the 1D B-p pass compares against
// bin edges past the 80 it initialises; here the comparison stops at the
// last initialised edge, which changes no printed value.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "gs_dropin.hpp"

namespace gsdrop {

namespace {

// Algorithm/include/ResultConstants.hpp:19-35
const double kM0 = 1.67261e-27;
const double kQ = 1.60219e-19;
const double kC = 2.99793e8;
const double kT0w = kM0 * kC * kC;
const double kT0 = kM0 * kC * kC / (kQ * 1e9);
const int kNT = 80;
const double kTminLog = 0.01, kTmaxLog = 200.0;
const double kTr = 0.938;
// SolarProp energies with a leading 0 (ResultConstants.hpp:39-44): index k
// (1..30) is 0.01 * 1.5^(k-1) GeV as printed to six digits.
const double kSPbins[32] = {0, 0.01, 0.015, 0.0225, 0.03375, 0.050625, 0.0759375, 0.113906,
                            0.170859, 0.256289, 0.384434, 0.57665, 0.864976, 1.29746, 1.9462,
                            2.91929, 4.37894, 6.56841, 9.85261, 14.7789, 22.1684, 33.2526,
                            49.8789, 74.8183, 112.227, 168.341, 252.512, 378.768, 568.151,
                            852.227, 1278.34, 0};
const double kUlyssesBins[4] = {0, 0.125, 0.250, 2.0};
// AMS-02 proton energy bin edges [GeV] (ResultConstants.hpp:47-61).
const double kAmsBins[72] = {
    4.924000e-01, 6.207000e-01, 7.637000e-01, 9.252000e-01, 1.105000e+00, 1.303000e+00,
    1.523000e+00, 1.765000e+00, 2.034000e+00, 2.329000e+00, 2.652000e+00, 3.005000e+00,
    3.390000e+00, 3.810000e+00, 4.272000e+00, 4.774000e+00, 5.317000e+00, 5.906000e+00,
    6.546000e+00, 7.236000e+00, 7.981000e+00, 8.787000e+00, 9.653000e+00, 1.060000e+01,
    1.160000e+01, 1.264000e+01, 1.379000e+01, 1.504000e+01, 1.639000e+01, 1.784000e+01,
    1.938000e+01, 2.103000e+01, 2.283000e+01, 2.478000e+01, 2.683000e+01, 2.903000e+01,
    3.138000e+01, 3.387000e+01, 3.657000e+01, 3.947000e+01, 4.257000e+01, 4.587000e+01,
    4.942000e+01, 5.322000e+01, 5.727000e+01, 6.162000e+01, 6.632000e+01, 7.137000e+01,
    7.677000e+01, 8.257000e+01, 8.882000e+01, 9.557000e+01, 1.031000e+02, 1.111000e+02,
    1.196000e+02, 1.291000e+02, 1.401000e+02, 1.526000e+02, 1.666000e+02, 1.826000e+02,
    2.006000e+02, 2.211000e+02, 2.451000e+02, 2.741000e+02, 3.096000e+02, 3.536000e+02,
    4.091000e+02, 4.821000e+02, 5.831000e+02, 7.316000e+02, 9.751000e+02, 1.464000e+03};

enum Kind { LOG = 1, LIN_1E2 = 2, LIN_1E3 = 3, LIN_4E2 = 4, SOLARPROP = 5, ULYSSES = 6, AMS = 7 };

double momentum_si(double T) {
    const double Tkinw = T * 1e9 * kQ;
    const double Rig = std::sqrt(Tkinw * (Tkinw + (2 * kT0w))) / kQ;
    return Rig * kQ / kC;
}

// ResultsUtils::writeSpectrum and its helpers.
void write_spectrum(std::string name, int size, int portion, bool csv, const double *count,
                    const double *value, Kind kind) {
    name += csv ? ".csv" : ".dat";
    FILE *f = std::fopen(name.c_str(), "w");
    if (!f) return;
    const char *s = csv ? "," : " ";
    if (csv) {
        if (kind == LOG) std::fprintf(f, "%s,%s\n", "binc", "spelog/binw");
        else std::fprintf(f, "%s,%s,%s,%s\n", "Tkin", "spe", "speAvg", "speCount");
    }
    const std::string f_log = std::string("%.14E") + s + "%.14E\n";
    const std::string f_four = std::string("%3.4f") + s + "%3.4f" + s + "%3.4f" + s + "%3.4f\n";
    const std::string f_lin = std::string("%3.4f") + s + "%.14E" + s + "%.14E\n";
    auto avg = [&](int i) { return count[i] != 0.0 ? value[i] / count[i] : 0.0; };
    if (kind == LOG) {
        double binb[25], binw[24], binc[24];
        for (int i = 0; i < 25; i++) binb[i] = std::exp((((i + 1) / 5.0) - 3.0) * std::log(10));
        for (int i = 0; i < 24; i++) {
            binw[i] = binb[i + 1] - binb[i];
            binc[i] = (binb[i + 1] + binb[i]) / 2.0;
        }
        for (int i = 0; i < 24; i++) std::fprintf(f, f_log.c_str(), binc[i], value[i] / binw[i]);
    } else if (kind == SOLARPROP) {
        for (int i = 1; i < 30; i++)
            std::fprintf(f, f_four.c_str(), (double)(0.01f * std::pow(1.5f, (float)(i - 1))), value[i], avg(i), count[i]);
    } else if (kind == ULYSSES) {
        for (int i = 1; i < 3; i++) std::fprintf(f, f_four.c_str(), kUlyssesBins[i], value[i], avg(i), count[i]);
    } else if (kind == AMS) {
        for (int i = 1; i < 71; i++) std::fprintf(f, f_four.c_str(), kAmsBins[i], value[i], avg(i), count[i]);
    } else {
        for (int i = 0; i < size; i++)
            std::fprintf(f, f_lin.c_str(), (double)((float)(i + 1) / portion), value[i], avg(i));
    }
    std::fclose(f);
}

// Lines of the file, then rewound (ResultsUtils::countLines).
long count_lines(FILE *f) {
    char buf[1024];
    long n = 0;
    while (std::fgets(buf, sizeof buf, f)) ++n;
    std::rewind(f);
    return n;
}

// The linear spectra shared by the 1D passes.
struct Linear {
    std::vector<double> s1e2 = std::vector<double>(100), n1e2 = std::vector<double>(100);
    std::vector<double> s4e2 = std::vector<double>(400), n4e2 = std::vector<double>(400);
    std::vector<double> s1e3 = std::vector<double>(1000), n1e3 = std::vector<double>(1000);
    void add(double Tkin, double w) {
        const int e2 = (int)std::trunc(Tkin);
        if (e2 > 0 && e2 < 101) { s1e2[e2 - 1] += w; n1e2[e2 - 1] += 1; }
        const int e3 = (int)std::trunc((Tkin + 0.05) * 10);
        if (e3 > 0 && e3 < 1001) { s1e3[e3 - 1] += w; n1e3[e3 - 1] += 1; }
        const int e4 = (int)std::trunc((Tkin + 0.125) * 4);
        if (e4 > 0 && e4 < 401) { s4e2[e4 - 1] += w; n4e2[e4 - 1] += 1; }
    }
    void write(bool csv, const double *spelog) {
        write_spectrum("output_1e3bin", 1000, 10, csv, n1e3.data(), s1e3.data(), LIN_1E3);
        write_spectrum("output_1e2bin", 100, 1, csv, n1e2.data(), s1e2.data(), LIN_1E2);
        write_spectrum("output_logbin", 100, 1, csv, nullptr, spelog, LOG);
        write_spectrum("output_4e2bin", 400, 4, csv, n4e2.data(), s4e2.data(), LIN_4E2);
    }
};

FILE *open_log(ParamsCarrier *pc, long *records) {
    FILE *f = std::fopen(pc->getString("pathToLogFile", "log.dat").c_str(), "r");
    if (!f) {
        std::fprintf(stderr, "[error] Could not open log file.\n");
        return nullptr;
    }
    *records = count_lines(f) - 1;
    if (*records < 0) std::fprintf(stderr, "[info] No trajectory found in log file.\n");
    return f;
}

}  // namespace

bool results_1d_bp(ParamsCarrier *pc) {
    long n = 0;
    FILE *in = open_log(pc, &n);
    if (!in) return false;
    const double dlT = std::log10(kTmaxLog / kTminLog) / kNT;
    const double X = std::log10(kTmaxLog);
    double binb[kNT];
    for (int i = 0; i < kNT; i++) binb[i] = std::exp((X - (i * dlT)) * std::log(10.0));
    std::vector<double> spelog(300, 0.0);
    Linear lin;
    for (long i = 0; i < n; i++) {
        double Tb, T1, r, w;   // boundary energy, 1 AU energy, r, printed weight
        if (std::fscanf(in, "%lf %lf %lf %lf\n", &Tb, &T1, &r, &w) == EOF) {
            std::fprintf(stderr, "[error] Could not read from log.dat file.\n");
            std::fclose(in);
            return false;
        }
        const double p = momentum_si(Tb);
        const double p1AU = momentum_si(T1);
        w = (kM0 * kM0 * kC * kC * kC * kC) + (p * p * kC * kC);
        w = std::exp(-1.85 * std::log(w)) / p;
        w = w * p1AU * p1AU;
        for (int ii = 0; ii + 1 < kNT; ii++)
            if ((T1 > binb[ii + 1]) && (T1 < binb[ii])) spelog[ii + 1] += w;
        lin.add(T1, w);
    }
    std::fclose(in);
    lin.write(pc->getInt("csv", 0) != 0, spelog.data());
    return true;
}

bool results_1d_fp(ParamsCarrier *pc) {
    long n = 0;
    FILE *in = open_log(pc, &n);
    if (!in) return false;
    double binb[25];
    for (int i = 0; i < 25; i++) binb[i] = std::exp((((i + 1) / 5.0) - 3.0) * std::log(10));
    std::vector<double> spelog(24, 0.0);
    Linear lin;
    for (long i = 0; i < n; i++) {
        double p100AU, p1AU, r, w0, sumac;
        if (std::fscanf(in, " %lf  %lf  %lf  %lf %lf \n", &p100AU, &p1AU, &r, &w0, &sumac) == EOF) {
            std::fprintf(stderr, "[error] Could not read from log.dat file.\n");
            std::fclose(in);
            return false;
        }
        const double Rig = p1AU * kC / kQ;
        double Tkin = std::sqrt((kT0 * kT0 * kQ * kQ * 1e9 * 1e9) + (kQ * kQ * Rig * Rig)) - (kT0 * kQ * 1e9);
        Tkin = Tkin / (kQ * 1e9);
        w0 = (kM0 * kM0 * kC * kC * kC * kC) + (p100AU * p100AU * kC * kC);
        w0 = std::exp(-1.85 * std::log(w0)) / p100AU;
        w0 = w0 / 1e45;
        double w = w0 * p1AU * p1AU * std::exp(sumac);
        if (r < 0.3) w = 0.0;
        for (int k = 0; k < 24; k++)
            if (Tkin > binb[k] && Tkin < binb[k + 1]) spelog[k] += w;
        lin.add(Tkin, w);
    }
    std::fclose(in);
    lin.write(pc->getInt("csv", 0) != 0, spelog.data());
    return true;
}

bool results_2d(ParamsCarrier *pc) {
    long n = 0;
    FILE *in = open_log(pc, &n);
    if (!in) return false;
    double speSP[31] = {0}, speJGR[31] = {0}, speN[31] = {0}, speBurger[31] = {0};
    double uly[4] = {0}, ulyN[4] = {0}, ams[72] = {0}, amsN[72] = {0};
    for (long i = 0; i < n; i++) {
        double Tinj, Tkin, r, w, thetainj, theta;
        if (std::fscanf(in, "%lf %lf %lf %lf %lf %lf\n", &Tinj, &Tkin, &r, &w, &thetainj, &theta) == EOF) {
            std::fprintf(stderr, "[error] Could not read from log.dat file.\n");
            std::fclose(in);
            return false;
        }
        const double p = momentum_si(Tkin);
        const double p1AU = momentum_si(Tinj);
        // Webber & Higbie LIS, evaluated at the 1 AU energy as the reference does.
        const double tem5 = 21.1 * std::exp(-2.8 * std::log(Tinj));
        const double tem6 = 1.0 + (5.85 * std::exp(-1.22 * std::log(Tinj))) + (1.18 * std::exp(-2.54 * std::log(Tinj)));
        w = (tem5 / tem6) / (p * p) * p1AU * p1AU;
        // Usoskin et al. 2005 LIS at the boundary energy.
        const double tt = Tkin + kTr, t2 = tt + kTr;
        const double beta = std::sqrt(Tkin * t2) / tt;
        double tem = std::exp(-3.93 * std::log((Tkin + 0.67) / 1.67));
        double jgr = 2.7e3 * std::exp(1.12 * std::log(Tkin)) * tem / (beta * beta);
        const double wJGR = jgr / (p * p) * p1AU * p1AU;
        // Burger et al. 2000 LIS at the boundary energy.
        tem = std::sqrt((Tkin * Tkin) + (2.0 * Tkin * 0.938));
        double jb = 1.9 * 1.0e4 * std::exp(-2.78 * std::log(tem));
        jb = jb / (1 + 0.4866 * std::exp(-2.51 * std::log(tem)));
        const double wB = jb / (p * p) * p1AU * p1AU;
        for (int ii = 1; ii < 30; ii++)
            if ((Tinj > kSPbins[ii] * 0.99) && (Tinj < kSPbins[ii] * 1.01)) {
                speSP[ii] += w;
                speJGR[ii] += wJGR;
                speBurger[ii] += wB;
                speN[ii]++;
            }
        for (int ii = 1; ii < 3; ii++)
            if ((Tinj > kUlyssesBins[ii]) && (Tinj < kUlyssesBins[ii + 1])) { uly[ii] += wB; ulyN[ii]++; }
        for (int ii = 1; ii < 71; ii++)
            if ((Tinj > kAmsBins[ii]) && (Tinj < kAmsBins[ii + 1])) { ams[ii] += wB; amsN[ii]++; }
    }
    std::fclose(in);
    const bool csv = pc->getInt("csv", 0) != 0;
    write_spectrum("Usoskin_2005", 0, 1, csv, speN, speJGR, SOLARPROP);
    write_spectrum("Weber", 0, 1, csv, speN, speSP, SOLARPROP);
    write_spectrum("Burger", 0, 1, csv, speN, speBurger, SOLARPROP);
    write_spectrum("Ulysses", 0, 1, csv, ulyN, uly, ULYSSES);
    write_spectrum("Ams", 0, 1, csv, amsN, ams, AMS);
    return true;
}

}  // namespace gsdrop
