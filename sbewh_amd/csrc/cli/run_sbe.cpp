// sbewh_run: the reference's command line SBE run over libsbewh_hip.so.
//
// Reads the reference's input decks (code/src/io/ParameterIO.cpp syntax, the
// parameters of code/src/io/*Parameter.cpp), builds the tight-binding model
// from the Wannier90 files (code/src/io/TightBindingParameter.cpp:111-148),
// applies the Nf / Nk rules (code/src/Parameter.cpp:4-56), computes the Fermi
// level and ground state, propagates every pulse configuration with the full
// observer and saves the results like code/src/io/save.cpp (npz by default,
// text files with Output.saveAsNpz = 0).  Multi-run files and their resume
// semantics follow code/src/main/runSBE.cpp:19-72,205-216: each configuration
// holds a lock file ".<name>" in the run directory, written with '1' when its
// results are saved; General.finishMultiRuns = 1 re-reads the arguments of the
// last (or the given) run directory and runs only the configurations that are
// neither finished nor running.
//
//   sbewh_run [input files] [-Group.name=value ...]      (default input: input.txt)
//   sbewh_run -par                                        (list the parameters)
#include <fcntl.h>
#include <sys/file.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "npz.hpp"
#include "params.hpp"
#include "sbewh_hip.h"
#include "sbewh_w90.h"

namespace fs = std::filesystem;
using cli::ParamError;
using cli::Registry;

namespace {

// code/src/util/unitConversion.h
constexpr double FS = 2.4188843266049513e-2, EV = 27.211386245771678, V_NM = 514.22067475617267,
                 NM = 5.291772109060855e-2, KELVIN = 315775.0248015561, C_AU = 137.035999177;

struct RegionSpec {
    std::string name;
    bool moving;
    double fwhm;
    unsigned dim[3];
    double k[3], dk[3][3];
};

struct Resolved {
    int dim = 3;
    std::string run_mode;
    sbe_grid grid{};
    sbe_params prm{};
    int out_count = 500;
    double smear = 0;
    bool store_density = false;
    int density_stride = 1;
    std::vector<RegionSpec> regions;
    std::vector<sbe_subpulse> pulse;                       // the pulse of the input files
    std::map<std::string, std::vector<sbe_subpulse>> configs;  // multi-run configurations (name order)
    std::string seed, multi_run;
};

[[noreturn]] void fail(const std::string &m) { throw ParamError{m}; }

// code/src/io/PulseParameter.cpp:121-190
std::vector<sbe_subpulse> build_pulse(const Registry &r, int dim)
{
    auto set = [&](const std::string &n, int i) { return r.is_set("Pulse", n + std::to_string(i)); };
    auto get = [&](const std::string &n, int i) -> std::string {
        if (i >= 0 && set(n, i)) return r.str("Pulse", n + std::to_string(i));
        return r.str("Pulse", n);
    };
    auto num = [&](const std::string &n, int i) {
        const std::string v = get(n, i);
        double x;
        if (std::sscanf(v.c_str(), "%lg", &x) != 1)
            fail("Pulse." + n + (i >= 0 ? std::to_string(i) : std::string()) + " is not set");
        return x;
    };
    const char *names[] = {"pol", "type", "Emax", "lambda", "cep", "FWHM", "tCentral", "tStart", "tEnd", "cyclesOn",
                           "risingCycles"};
    {
        std::vector<double> pd = r.vec("Pulse", "pol");
        if (!pd.empty() && pd.size() != 3) fail("Default polarization must have 3 entries");
        double a, b;
        if (std::sscanf(r.str("Pulse", "tStart").c_str(), "%lg", &a) == 1 &&
            std::sscanf(r.str("Pulse", "tEnd").c_str(), "%lg", &b) == 1 && a >= b)
            fail("Simulation start time has to be before simulation end time");
    }
    std::vector<int> ids;
    for (int i = 0; i < 10; i++)
        for (const char *n : names)
            if (set(n, i)) {
                ids.push_back(i);
                break;
            }
    if (ids.empty()) ids.push_back(-1);
    std::vector<sbe_subpulse> out;
    for (int i : ids) {
        sbe_subpulse s{};
        std::vector<double> pol;
        {
            std::string v = get("pol", i);
            char *copy = strdup(v.c_str());
            for (char *t = std::strtok(copy, ", "); t; t = std::strtok(nullptr, ", ")) pol.push_back(std::atof(t));
            std::free(copy);
        }
        if (pol.size() != 3) fail("Polarization for pulse " + std::to_string(i) + " must have exactly 3 entries");
        for (int d = dim; d < 3; d++)
            if (std::fabs(pol[d]) > 1e-10)
                std::fprintf(stderr, "warning: polarization of subpulse %d has a non-zero component in direction %d\n", i,
                             d + 1);
        const std::string type = get("type", i);
        if (type == "sin2") s.type = SBE_PULSE_SIN2;
        else if (type == "gauss") s.type = SBE_PULSE_GAUSS;
        else if (type == "pureGauss") s.type = SBE_PULSE_PURE_GAUSS;
        else if (type == "sin2Ramp") s.type = SBE_PULSE_SIN2_RAMP;
        else fail("Unrecognized pulse envelope type '" + type + "' for pulse " + std::to_string(i));
        s.E0 = num("Emax", i) / V_NM;
        s.cep = num("cep", i);
        s.center = num("tCentral", i) / FS;
        s.cycles_on = (int)num("cyclesOn", i);
        s.rising_cycles = num("risingCycles", i);
        if (s.type != SBE_PULSE_PURE_GAUSS || !get("lambda", i).empty()) s.omega = 2 * M_PI * C_AU / (num("lambda", i) / NM);
        if (s.type != SBE_PULSE_SIN2_RAMP) {
            s.fwhm = num("FWHM", i) / FS;
            s.t_start_fwhm = num("tStart", i);
            s.t_end_fwhm = num("tEnd", i);
            if (s.t_start_fwhm >= s.t_end_fwhm) fail("tStart must be before tEnd (pulse " + std::to_string(i) + ")");
        }
        for (int d = 0; d < 3; d++) s.pol[d] = pol[d];
        out.push_back(s);
    }
    return out;
}

// code/src/io/KspaceRegionParameter.cpp:78-164
RegionSpec parse_region(const std::string &name, const std::string &values)
{
    char mode = 0, range = 'E';
    double k[12];
    unsigned count[3] = {1, 1, 1};
    double fwhm = 0;
    int dim = -1;
    const char *v = values.c_str();
    if (std::sscanf(v, "%c %lg %lg %lg %lg %lg %lg %lg %lg %lg %lg %lg %lg %u %u %u %c %lg", &mode, k, k + 1, k + 2,
                    k + 3, k + 4, k + 5, k + 6, k + 7, k + 8, k + 9, k + 10, k + 11, count, count + 1, count + 2, &range,
                    &fwhm) == 18)
        dim = 3;
    if (dim < 0 && std::sscanf(v, "%c %lg %lg %lg %lg %lg %lg %lg %lg %lg %u %u %c %lg", &mode, k, k + 1, k + 2, k + 3,
                               k + 4, k + 5, k + 6, k + 7, k + 8, count, count + 1, &range, &fwhm) == 14)
        dim = 2;
    if (dim < 0 && std::sscanf(v, "%c %lg %lg %lg %lg %lg %lg %u %c %lg", &mode, k, k + 1, k + 2, k + 3, k + 4, k + 5,
                               count, &range, &fwhm) == 10)
        dim = 1;
    if (dim < 0 && std::sscanf(v, "%c %lg %lg %lg %lg", &mode, k, k + 1, k + 2, &fwhm) == 5) dim = 0;
    if (dim < 0) fail(name + ": could not properly scan KspaceRegionParamters '" + values + "'");
    if (mode != 'F' && mode != 'M') fail(name + ": mode must be either 'F' (fixed frame) or 'M' (moving frame)");
    if (dim > 0 && range != 'I' && range != 'E') fail(name + ": range type must be either 'I' (inclusive) or 'E' (exclusive)");
    for (int i = 0; i < (dim + 1) * 3; i++)
        if (k[i] < -1e-10 || k[i] > 1 + 1e-10) fail(name + ": k must be in range [0, 1]");
    if (fwhm <= 0 || fwhm >= 1) fail(name + ": full width half maximum must be bigger 0 and smaller 1");
    RegionSpec r{};
    r.name = name;
    r.moving = mode == 'M';
    r.fwhm = fwhm;
    for (int d = 0; d < 3; d++) r.k[d] = k[d];
    for (int i = 0; i < 3; i++) {
        r.dim[i] = 1;
        for (int d = 0; d < 3; d++) r.dk[i][d] = 0;
    }
    for (int i = 0; i < dim; i++) {
        if (count[i] < 2) fail(name + ": specified sampling dimensions must all be bigger than 1");
        const unsigned neff = range == 'E' ? count[i] : count[i] - 1;
        r.dim[i] = count[i];
        for (int d = 0; d < 3; d++) r.dk[i][d] = (k[3 * i + 3 + d] - k[d]) / (double)neff;
    }
    return r;
}

// code/src/Parameter.cpp:4-56 with the model's cell extent Nw
void validate_k(Resolved &R, const int Nw[3], unsigned max_prime, bool allow_adapt, int Nf[3])
{
    int maxNf = 0;
    for (int d = 0; d < R.dim; d++) Nf[d] = std::max(Nf[d], Nw[d]);
    for (int d = 0; d < 3; d++) maxNf = std::max(maxNf, Nf[d]);
    if ((unsigned)maxNf > max_prime) {
        std::vector<unsigned> primes{2};
        for (unsigned s = 3; s <= max_prime; s++) {
            unsigned d = 2;
            while (d < s && s % d) d++;
            if (d == s) primes.push_back(s);
        }
        for (int d = 0; d < R.dim; d++) {
            unsigned N = Nf[d] - 1, Nc;
            do {
                Nc = ++N;
                for (unsigned p : primes)
                    while (Nc % p == 0) Nc /= p;
            } while (Nc != 1);
            Nf[d] = (int)N;
        }
    }
    for (int d = 0; d < R.dim; d++) {
        if (Nf[d] == 1) continue;
        if (R.grid.Nk[d] < Nf[d])
            fail("In k-dimension " + std::to_string(d + 1) + ": Nk (" + std::to_string(R.grid.Nk[d]) +
                 ") is smaller than Nf (" + std::to_string(Nf[d]) + ")");
        const int rem = R.grid.Nk[d] % Nf[d];
        if (rem) {
            if (!allow_adapt)
                fail("In k-dimension " + std::to_string(d + 1) + ": Nk (" + std::to_string(R.grid.Nk[d]) +
                     ") is not a multiple of Nf (" + std::to_string(Nf[d]) + ")");
            R.grid.Nk[d] -= rem;
            if (rem >= Nf[d] / 2) R.grid.Nk[d] += Nf[d];
        }
    }
}

std::string in_dir(const std::string &fixed_dir, const std::string &p)
{
    if (fixed_dir.empty() || p.empty()) return p;
    return (fs::path(fixed_dir) / fs::path(p).filename()).string();
}

Resolved resolve(Registry &r, const std::string &fixed_dir)
{
    Resolved R;
    R.run_mode = r.str("General", "runMode");
    R.dim = (int)r.integer("General", "dimensionality");
    if (R.dim < 1 || R.dim > 3) fail("dimensionality must be 1, 2 or 3");
    R.grid.dim = R.dim;
    {
        std::vector<double> nk = r.vec("Propagation", "Nk");
        if (nk.size() != 1 && (int)nk.size() < R.dim) fail("Wrong number of dimensions in Nk");
        for (int d = 0; d < 3; d++) R.grid.Nk[d] = d < R.dim ? (int)(nk.size() == 1 ? nk[0] : nk[d]) : 1;
        for (size_t d = R.dim; d < nk.size() && nk.size() > 1; d++)
            if (nk[d] != 1) fail(std::to_string(d + 1) + ". component of Nk must be 1 or omitted");
        std::vector<double> nf = r.vec("FourierTransform", "Nf");
        for (int d = 0; d < 3; d++) R.grid.Nf[d] = d < R.dim ? (int)(nf.size() == 1 ? nf[0] : d < (int)nf.size() ? nf[d] : 1) : 1;
        std::vector<double> ks = r.vec("Propagation", "kGlobalShift");
        if (ks.size() > 3) fail("kGlobalShift must have at most three dimensions");
        for (size_t d = 0; d < ks.size(); d++) {
            if (ks[d] < -1e-10 || ks[d] > 1 + 1e-10) fail("kGlobalShift entries must be in range [0, 1]");
            R.grid.k_global_shift[d] = ks[d];
        }
    }
    R.out_count = (int)r.integer("Propagation", "outputCount");
    R.prm.relaxation_F2 = 1 / (r.num("Propagation", "T2") / FS);
    R.prm.soothing_width = r.num("Propagation", "soothingWidth") / EV;
    R.smear = r.is_set("Propagation", "occupationSmearingWidth") ? r.num("Propagation", "occupationSmearingWidth") / EV
                                                                  : R.prm.soothing_width;
    R.prm.gs_temp = r.num("TightBinding", "temperature") / KELVIN;
    R.prm.eps_abs = r.num("Propagation", "epsAbs");
    R.prm.eps_rel = r.num("Propagation", "epsRel");
    if (R.prm.eps_abs < 0 || R.prm.eps_abs > 1e-2 || R.prm.eps_rel < 0 || R.prm.eps_rel > 1e-2)
        fail("epsAbs and epsRel must be in [0, 1e-2]");
    R.prm.sweep_count = (int)r.integer("Diagonalization", "maxSweepCount");
    R.store_density = r.flag("KspaceRegions", "storeDensityMatrix");
    R.density_stride = (int)r.integer("KspaceRegions", "densityOutputStride");
    if (R.store_density && R.density_stride <= 0)
        fail("densityOutputStride must be bigger than zero if storeDensityMatrix is enabled");
    // the whole zone first, then the named regions in name order
    RegionSpec zone{};
    zone.name = "expectationValues";
    zone.moving = true;
    zone.fwhm = INFINITY;
    for (int i = 0; i < 3; i++) zone.dim[i] = 1;
    R.regions.push_back(zone);
    for (const auto &[name, value] : r.custom["KspaceRegions"]) R.regions.push_back(parse_region(name, value));
    R.seed = in_dir(fixed_dir, r.str("TightBinding", "wannierSeed"));
    R.pulse = build_pulse(r, R.dim);
    R.multi_run = in_dir(fixed_dir, r.str("Pulse", "multiRunFile"));
    if (!R.multi_run.empty() && fs::exists(R.multi_run)) {
        // configurations [name] of Pulse parameters over the input files' values
        auto base = r.snapshot();
        std::string pending;
        auto close_config = [&](const std::string &next) {
            if (!pending.empty()) R.configs[pending] = build_pulse(r, R.dim);
            r.restore(base);
            pending = next;
        };
        FILE *f = std::fopen(R.multi_run.c_str(), "r");
        if (!f) fail("Could not parse multiconfiguration file '" + R.multi_run + "'");
        try {
            r.parse_stream(f, "Pulse", &close_config);
        } catch (...) {
            std::fclose(f);
            throw;
        }
        std::fclose(f);
        close_config("");
    }
    return R;
}

std::vector<size_t> squeeze(std::initializer_list<size_t> l)
{
    std::vector<size_t> s;
    for (size_t v : l)
        if (v > 1) s.push_back(v);
    return s;
}

struct Results {
    int count = 0, dcount = 0, R = 0, n = 0;
    std::vector<double> t, E, A, ws, j, oh, ow, mn, mx, wan;
    std::vector<int> didx;
};

// code/src/io/save.cpp:182-211
bool save_npz(const std::string &path, const Results &r, const std::vector<RegionSpec> &regs, const double lat[9],
              bool density)
{
    cli::Npz z(path);
    if (!z.ok()) return false;
    const size_t Nt = r.count, n = r.n, Nd = r.dcount;
    z.add("t", r.t.data(), {Nt});
    z.add("pulse/E", r.E.data(), {Nt, 3});
    z.add("pulse/A", r.A.data(), {Nt, 3});
    z.add("lattice", lat, {3, 3});
    z.add("occHamMin", r.mn.data(), {Nt, n});
    z.add("occHamMax", r.mx.data(), {Nt, n});
    if (density) {
        std::vector<uint32_t> di(r.didx.begin(), r.didx.end());
        z.add("tDensIndices", di.data(), {Nd});
    }
    size_t off = 0;
    for (const auto &g : regs) {
        const size_t d0 = g.dim[0], d1 = g.dim[1], d2 = g.dim[2], ns = d0 * d1 * d2;
        std::vector<double> w(Nt * ns), j(Nt * ns * 3), oh(Nt * ns * n), ow(Nt * ns * n);
        for (size_t ti = 0; ti < Nt; ti++)
            for (size_t s = 0; s < ns; s++) {
                const size_t src = ti * r.R + off + s, dst = ti * ns + s;
                w[dst] = r.ws[src];
                for (int d = 0; d < 3; d++) j[3 * dst + d] = r.j[3 * src + d];
                for (size_t b = 0; b < n; b++) {
                    oh[dst * n + b] = r.oh[src * n + b];
                    ow[dst * n + b] = r.ow[src * n + b];
                }
            }
        z.add(g.name + "/weight", w.data(), squeeze({Nt, d0, d1, d2}));
        z.add(g.name + "/j", j.data(), squeeze({Nt, d0, d1, d2, 3}));
        z.add(g.name + "/occW", ow.data(), squeeze({Nt, d0, d1, d2, n}));
        z.add(g.name + "/occH", oh.data(), squeeze({Nt, d0, d1, d2, n}));
        if (density) {
            std::vector<std::complex<double>> wan(Nd * ns * n * n);
            for (size_t ti = 0; ti < Nd; ti++)
                for (size_t s = 0; s < ns; s++)
                    for (size_t e = 0; e < n * n; e++) {
                        const size_t src = ((ti * r.R + off + s) * n * n + e) * 2;
                        wan[(ti * ns + s) * n * n + e] = std::complex<double>(r.wan[src], r.wan[src + 1]);
                    }
            z.add(g.name + "/wan", wan.data(), squeeze({Nd, d0, d1, d2, n, n}));
        }
        off += ns;
    }
    return z.close();
}

// code/src/io/save.cpp:29-140 (text files of one configuration)
bool save_txt(const fs::path &dir, const Results &r, const std::vector<RegionSpec> &regs, bool density)
{
    const int n = r.n;
    FILE *f = std::fopen((dir / "pulse.txt").c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "# all in atomic units\n# t\tEx\tEy\tEz\tAx\tAy\tAz\n");
    for (int i = 0; i < r.count; i++) {
        std::fprintf(f, "%.5lg", r.t[i]);
        for (int d = 0; d < 3; d++) std::fprintf(f, "\t%.10lg", r.E[3 * i + d]);
        for (int d = 0; d < 3; d++) std::fprintf(f, "\t%.10lg", r.A[3 * i + d]);
        std::fprintf(f, "\n");
    }
    if (std::fclose(f)) return false;
    f = std::fopen((dir / "occupationLimits.txt").c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "# minimum and maximum occupation over k-points\n# t");
    for (int b = 0; b < n; b++) std::fprintf(f, "\toccHamMin%d\toccHamMax%d", b, b);
    std::fprintf(f, "\n");
    for (int i = 0; i < r.count; i++) {
        std::fprintf(f, "%lg", r.t[i]);
        for (int b = 0; b < n; b++) std::fprintf(f, "\t%.10lg\t%.10lg", r.mn[(size_t)n * i + b], r.mx[(size_t)n * i + b]);
        std::fprintf(f, "\n");
    }
    if (std::fclose(f)) return false;
    size_t off = 0;
    for (const auto &g : regs) {
        unsigned si = 0;
        for (unsigned a = 0; a < g.dim[0]; a++)
            for (unsigned b = 0; b < g.dim[1]; b++)
                for (unsigned c = 0; c < g.dim[2]; c++, si++) {
                    std::string name = g.name;
                    if (g.dim[0] > 1) name += "_" + std::to_string(a);
                    if (g.dim[1] > 1) name += "_" + std::to_string(b);
                    if (g.dim[2] > 1) name += "_" + std::to_string(c);
                    FILE *e = std::fopen((dir / (name + ".txt")).c_str(), "w");
                    if (!e) return false;
                    std::fprintf(e, "# all in atomic units\n# t\tweight\tJx\tJy\tJz\t");
                    for (int q = 0; q < n; q++) std::fprintf(e, "\toccW%d", q);
                    for (int q = 0; q < n; q++) std::fprintf(e, "\toccH%d", q);
                    std::fprintf(e, "\n");
                    for (int ti = 0; ti < r.count; ti++) {
                        const size_t s = (size_t)ti * r.R + off + si;
                        std::fprintf(e, "%lg\t%lg", r.t[ti], r.ws[s]);
                        for (int d = 0; d < 3; d++) std::fprintf(e, "\t%lg", r.j[3 * s + d]);
                        for (int q = 0; q < n; q++) std::fprintf(e, "\t%.10lg", r.ow[s * n + q]);
                        for (int q = 0; q < n; q++) std::fprintf(e, "\t%.10lg", r.oh[s * n + q]);
                        std::fprintf(e, "\n");
                    }
                    if (std::fclose(e)) return false;
                    if (density) {
                        FILE *h = std::fopen((dir / ("rho_" + name + ".txt")).c_str(), "w");
                        if (!h) return false;
                        std::fprintf(h, "# interpolated density matrix, time in atomic units\n# t\tweight");
                        for (int p = 0; p < n; p++)
                            for (int q = 0; q < n; q++) std::fprintf(h, "\tRe_%d_%d\tIm_%d_%d", p, q, p, q);
                        std::fprintf(h, "\n");
                        for (int di = 0; di < r.dcount; di++) {
                            // the reference's weight column indexes the weights by the density record
                            const size_t sw = (size_t)di * r.R + off + si;
                            std::fprintf(h, "%lg\t%lg", r.t[r.didx[di]], r.ws[sw]);
                            for (int e2 = 0; e2 < n * n; e2++) {
                                const size_t src = (((size_t)di * r.R + off + si) * n * n + e2) * 2;
                                std::fprintf(h, "\t%.10lg\t%.10lg", r.wan[src], r.wan[src + 1]);
                            }
                            std::fprintf(h, "\n");
                        }
                        if (std::fclose(h)) return false;
                    }
                }
        off += g.dim[0] * g.dim[1] * g.dim[2];
    }
    return true;
}

// prepareMultiRun / finishMultiRun, code/src/main/runSBE.cpp:19-72: -1 when the
// configuration is finished or held by another process
int prepare_multi_run(const fs::path &dir, const std::string &name)
{
    const std::string lock = (dir / ("." + name)).string();
    int fd = open(lock.c_str(), O_WRONLY | O_CREAT | O_EXCL, S_IRUSR | S_IWUSR);
    if (fd == -1 && errno == EEXIST) {
        fd = open(lock.c_str(), O_WRONLY);
        if (fd == -1) return -1;
        if (flock(fd, LOCK_EX | LOCK_NB) && errno == EWOULDBLOCK) {
            close(fd);
            return -1;
        }
        if (lseek(fd, 0, SEEK_END) != 0) {
            close(fd);
            return -1;
        }
        return fd;
    }
    if (fd == -1) return -1;
    if (flock(fd, LOCK_EX | LOCK_NB) && errno == EWOULDBLOCK) {
        close(fd);
        return -1;
    }
    return fd;
}

void finish_multi_run(int fd)
{
    if (fd == -1) return;
    const char s = '1';
    if (write(fd, &s, 1) != 1) std::fprintf(stderr, "warning: finished flag could not be written\n");
    flock(fd, LOCK_UN);
    close(fd);
}

struct Lib {
    static void check(int rc, sbe_ctx *c, const char *what)
    {
        if (rc < 0) throw ParamError{std::string(what) + ": " + sbe_last_error(c)};
    }
};

Results collect(sbe_observer *obs, sbe_ctx *ctx, int R, int n)
{
    Results r;
    int dc = 0;
    r.count = sbe_observer_count(obs, &dc);
    r.dcount = dc;
    r.R = R;
    r.n = n;
    const size_t N = r.count;
    r.t.resize(N);
    r.E.resize(3 * N);
    r.A.resize(3 * N);
    r.ws.resize(N * R);
    r.j.resize(3 * N * R);
    r.oh.resize(N * R * n);
    r.ow.resize(N * R * n);
    r.mn.resize(N * n);
    r.mx.resize(N * n);
    r.wan.resize((size_t)dc * R * n * n * 2 + 1);
    r.didx.resize(dc + 1);
    sbe_observations o{r.t.data(),  r.E.data(),  r.A.data(),  r.ws.data(),  r.j.data(),  r.oh.data(),
                       r.ow.data(), r.mn.data(), r.mx.data(), r.wan.data(), r.didx.data()};
    Lib::check(sbe_observer_results(obs, &o), ctx, "sbe_observer_results");
    return r;
}

int run(int argc, char **argv)
{
    Registry reg;
    cli::register_sbe(reg);
    std::vector<std::string> args(argv + 1, argv + argc);
    if (args.size() == 2 && args[0] == "--write-test-npz") {
        // a fixed set of members, for checking the container against numpy.load
        const double t[3] = {0.0, 0.5, 1.25};
        double E[6];
        for (int i = 0; i < 6; i++) E[i] = 0.1 * i - 0.2;
        const std::complex<double> c[4] = {{1, -1}, {0.5, 2}, {-3, 0.25}, {0, 1e-300}};
        const uint32_t u[2] = {7, 4000000000u};
        cli::Npz z(args[1]);
        z.add("t", t, {3});
        z.add("pulse/E", E, {3, 2});
        z.add("region/wan", c, {2, 2});
        z.add("tDensIndices", u, {2});
        return z.close() ? 0 : 1;
    }
    for (const auto &a : args)
        if (a == "-par" || a == "--par") {
            for (const auto &[grp, ps] : reg.g)
                for (const auto &[name, p] : ps)
                    std::printf("%s.%s = %s\n", grp.c_str(), name.c_str(), p.value.c_str());
            return 0;
        } else if (a == "-h" || a == "--help" || a == "-help") {
            std::printf("usage: sbewh_run [input files] [-Group.name=value ...]\n"
                        "       sbewh_run -par   (parameters and defaults)\n");
            return 0;
        }
    auto parse_args = [&](Registry &r, const std::vector<std::string> &as) {
        bool any_file = false;
        for (const auto &a : as) any_file |= a[0] != '-';
        if (!any_file) r.parse_file("input.txt");
        for (const auto &a : as) {
            if (a[0] != '-') r.parse_file(a);
            else r.parse_option(a);
        }
    };
    parse_args(reg, args);
    std::string fixed_dir;
    fs::path save_dir;
    const bool resume = reg.flag("General", "finishMultiRuns");
    const bool disable = reg.flag("Output", "disableRunSaving");
    // the run directory (code/src/io/OutputParameter.cpp:17-70)
    {
        const fs::path base = reg.str("Output", "outputDirectory");
        std::string run = reg.str("Output", "run");
        if (!disable) {
            if (!run.empty()) {
                const bool exists = fs::is_directory(base / run);
                if (exists && !resume) fail("Run '" + run + "' is already existent in '" + base.string() + "'");
                if (!exists && resume) fail("Run '" + run + "' was not setup yet in '" + base.string() + "'");
            } else {
                int nr = 0;
                char name[16];
                if (fs::is_directory(base)) {
                    do std::snprintf(name, sizeof name, "%03d", ++nr);
                    while (fs::exists(base / name));
                    if (resume && --nr < 1) fail("output directory '" + base.string() + "' contains no automatically chosen sub run");
                } else {
                    if (resume) fail("output directory '" + base.string() + "' not yet existent");
                    nr = 1;
                }
                std::snprintf(name, sizeof name, "%03d", nr);
                run = name;
            }
            save_dir = base / run;
        }
    }
    if (resume) {
        // the arguments of the run being finished, with its copied input files
        std::ifstream in(save_dir / "args.txt");
        if (!in) fail("Could not read '" + (save_dir / "args.txt").string() + "'");
        std::string line;
        std::vector<std::string> old;
        std::getline(in, line);
        std::getline(in, line);
        while (std::getline(in, line))
            if (!line.empty()) old.push_back(line);
        Registry again;
        cli::register_sbe(again);
        fixed_dir = save_dir.string();
        std::vector<std::string> moved;
        for (const auto &a : old) moved.push_back(a[0] == '-' ? a : in_dir(fixed_dir, a));
        parse_args(again, moved);
        if (again.flag("General", "finishMultiRuns")) fail("Recursive call of finishMultiRuns is not allowed");
        reg = again;
    }
    Resolved R = resolve(reg, fixed_dir);
    if (R.run_mode == "Dry" || reg.flag("General", "dryRun")) {
        std::printf("{\"runMode\": \"%s\", \"dim\": %d, \"Nk\": [%d, %d, %d], \"Nf\": [%d, %d, %d], \"outputCount\": %d, "
                    "\"F2\": %.17g, \"soothingWidth\": %.17g, \"smearing\": %.17g, \"epsAbs\": %.17g, \"epsRel\": %.17g, "
                    "\"sweeps\": %d, \"kT\": %.17g, \"regions\": %zu, \"pulses\": %zu, \"configs\": %zu, \"seed\": \"%s\"}\n",
                    R.run_mode.c_str(), R.dim, R.grid.Nk[0], R.grid.Nk[1], R.grid.Nk[2], R.grid.Nf[0], R.grid.Nf[1],
                    R.grid.Nf[2], R.out_count, R.prm.relaxation_F2, R.prm.soothing_width, R.smear, R.prm.eps_abs,
                    R.prm.eps_rel, R.prm.sweep_count, R.prm.gs_temp, R.regions.size(), R.pulse.size(), R.configs.size(),
                    R.seed.c_str());
        for (const auto &g : R.regions)
            std::printf("region %s %c fwhm %.17g dims %u %u %u k %.17g %.17g %.17g\n", g.name.c_str(), g.moving ? 'M' : 'F',
                        g.fwhm, g.dim[0], g.dim[1], g.dim[2], g.k[0], g.k[1], g.k[2]);
        for (const auto &[name, p] : R.configs)
            std::printf("config %s: %zu sub-pulses, E0 %.17g omega %.17g cep %.17g\n", name.c_str(), p.size(), p[0].E0,
                        p[0].omega, p[0].cep);
        return 0;
    }
    if (R.run_mode != "SBE") fail("run mode '" + R.run_mode + "' is not provided by this build (SBE and Dry are)");
    if (R.seed.empty()) fail("No wannier seed set");
    // model (code/src/io/TightBindingParameter.cpp:54-148)
    sbe_w90_tb *tb = nullptr;
    sbe_w90_wsvec *ws = nullptr;
    sbe_w90_model *model = nullptr;
    if (sbe_w90_tb_read((R.seed + "_tb.dat").c_str(), &tb)) fail(std::string("Could not read '") + R.seed + "_tb.dat': " + sbe_w90_last_error());
    if (fs::exists(R.seed + "_wsvec.dat") && sbe_w90_wsvec_read((R.seed + "_wsvec.dat").c_str(), &ws))
        fail(std::string("Could not read '") + R.seed + "_wsvec.dat': " + sbe_w90_last_error());
    const int flags = (reg.flag("TightBinding", "symmetrizeHamiltonian") ? SBE_W90_SYMMETRIZE : 0) |
                      (reg.flag("TightBinding", "onlyRealMatrixElements") ? SBE_W90_ONLY_REAL : 0);
    if (sbe_w90_model_build(tb, ws, flags, R.dim, &model)) fail(std::string("model: ") + sbe_w90_last_error());
    sbe_model view{};
    sbe_w90_model_view(model, &view);
    const int n = view.num_wann;
    int Nw[3] = {1, 1, 1};
    {
        int mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
        for (int c = 0; c < view.num_cells; c++)
            for (int d = 0; d < 3; d++) {
                mn[d] = std::min(mn[d], view.cells[3 * c + d]);
                mx[d] = std::max(mx[d], view.cells[3 * c + d]);
            }
        for (int d = 0; d < 3; d++) Nw[d] = mx[d] - mn[d] + 1;
    }
    if ((unsigned)reg.integer("FourierTransform", "maxNfPrime") != 5) fail("FourierTransform.maxNfPrime other than 5 is not provided");
    validate_k(R, Nw, 5, reg.flag("Propagation", "allowNkAdaption"), R.grid.Nf);
    const bool custom_mu = reg.is_set("TightBinding", "fermiLevel");
    const bool below_gap = reg.flag("TightBinding", "occupiedBelowBandGap");
    const long occ = reg.integer("TightBinding", "occupiedBands");
    if (!custom_mu && !below_gap) {
        if (occ == 0) fail("Either occupiedBands > 0 or occupiedBelowBandGap must be set.");
        if (occ >= n) fail("occupiedBands >= numWann");
    }
    // output directory and saved input (code/src/io/save.cpp:4-27)
    if (!disable) {
        fs::create_directories(save_dir);
        if (!resume && reg.flag("Output", "saveInput")) {
            for (const auto &f : reg.files) fs::copy_file(f, save_dir / fs::path(f).filename(), fs::copy_options::overwrite_existing);
            FILE *fa = std::fopen((save_dir / "args.txt").c_str(), "w");
            std::fprintf(fa, "Git version: %s\nCommand line arguments (one per line):\n", sbe_version());
            for (const auto &a : args) std::fprintf(fa, "%s\n", a.c_str());
            std::fclose(fa);
            const bool wan = reg.is_set("Output", "saveWannierInput") ? reg.flag("Output", "saveWannierInput") : true;
            if (wan)
                for (const std::string &suffix : {std::string("_tb.dat"), std::string("_wsvec.dat")})
                    if (fs::exists(R.seed + suffix))
                        fs::copy_file(R.seed + suffix, save_dir / fs::path(R.seed + suffix).filename(),
                                      fs::copy_options::overwrite_existing);
            if (!R.multi_run.empty() && fs::exists(R.multi_run))
                fs::copy_file(R.multi_run, save_dir / fs::path(R.multi_run).filename(), fs::copy_options::overwrite_existing);
        }
    }
    sbe_ctx *ctx = nullptr;
    const std::vector<sbe_subpulse> &first = R.configs.empty() ? R.pulse : R.configs.begin()->second;
    Lib::check(sbe_create(&view, &R.grid, &R.prm, first.data(), (int)first.size(), &ctx), nullptr, "sbe_create");
    double mu = 0;
    if (custom_mu) mu = reg.num("TightBinding", "fermiLevel") / EV;
    else Lib::check(sbe_fermi_level(ctx, (int)occ, below_gap ? 1 : 0, &mu), ctx, "Fermi level");
    std::printf("Fermi level: %.10g eV, %ld k-points, %d bands\n", mu * EV, sbe_num_k(ctx), n);
    std::vector<double> ground(2 * sbe_state_size(ctx));
    Lib::check(sbe_initial_state(ctx, mu, ground.data()), ctx, "initial state");
    std::vector<sbe_region> samples;
    for (const auto &g : R.regions)
        for (unsigned a = 0; a < g.dim[0]; a++)
            for (unsigned b = 0; b < g.dim[1]; b++)
                for (unsigned c = 0; c < g.dim[2]; c++) {
                    sbe_region s{};
                    for (int d = 0; d < 3; d++) s.k[d] = g.k[d] + a * g.dk[0][d] + b * g.dk[1][d] + c * g.dk[2][d];
                    s.fwhm = g.fwhm;
                    s.moving_frame = g.moving ? 1 : 0;
                    samples.push_back(s);
                }
    sbe_observer *obs = nullptr;
    Lib::check(sbe_observer_create(ctx, samples.data(), (int)samples.size(), R.smear,
                                   R.store_density ? R.density_stride : 0, R.out_count, &obs),
               ctx, "observer");
    auto propagate = [&](const std::vector<sbe_subpulse> &pulse, const std::string &name) {
        Lib::check(sbe_set_pulses(ctx, pulse.data(), (int)pulse.size()), ctx, "pulse");
        Lib::check(sbe_observer_reset(obs), ctx, "observer reset");
        std::vector<double> st = ground;
        long rhs = 0;
        const auto t0 = std::chrono::steady_clock::now();
        const long steps = sbe_integrate_observed(ctx, st.data(), R.out_count, obs, nullptr, nullptr, nullptr, nullptr, &rhs);
        if (steps < 0) Lib::check((int)steps, ctx, "integration");
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("%s: %ld steps, %ld RHS evaluations, %.3f s\n", name.empty() ? "run" : name.c_str(), steps, rhs, secs);
        std::fflush(stdout);
        if (disable) return;
        Results res = collect(obs, ctx, (int)samples.size(), n);
        bool ok;
        if (reg.flag("Output", "saveAsNpz")) {
            ok = save_npz((save_dir / ((name.empty() ? std::string("data") : name) + ".npz")).string(), res, R.regions,
                          view.lattice, R.store_density);
        } else {
            fs::path d = name.empty() ? save_dir : save_dir / name;
            fs::create_directories(d);
            ok = save_txt(d, res, R.regions, R.store_density);
        }
        if (!ok) fail("results of '" + name + "' not written properly");
    };
    if (!R.configs.empty()) {
        for (const auto &[name, pulse] : R.configs) {
            const int fd = disable ? 0 : prepare_multi_run(save_dir, name);
            if (fd == -1) continue;
            std::printf("Configuration: %s\n", name.c_str());
            propagate(pulse, name);
            if (!disable) finish_multi_run(fd);
        }
    } else {
        propagate(R.pulse, "");
    }
    sbe_observer_destroy(obs);
    sbe_destroy(ctx);
    sbe_w90_model_free(model);
    sbe_w90_tb_free(tb);
    if (ws) sbe_w90_wsvec_free(ws);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return run(argc, argv);
    } catch (const ParamError &e) {
        std::fprintf(stderr, "error: %s\n", e.msg.c_str());
        return 1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
