// Input decks of the reference's command line programs: the parameter file
// syntax of code/src/io/ParameterIO.cpp:137-296 (sections [Group], lines
// "name = value", '#' comments, '\' line continuation, dotted Group.name,
// command line options -Group.name=value) and the parameters the SBE run mode
// reads, with the reference's names, defaults and value syntax
// (code/src/io/*Parameter.cpp).  Values keep the reference's sscanf
// semantics: "2 fs" reads as 2, booleans are true / false / integers, vectors
// split at ',' and ' '.
#pragma once
#include <cstdio>
#include <cstring>
#include <map>
#include <regex>
#include <string>
#include <vector>

namespace cli {

struct ParamError {
    std::string msg;
};

enum class Kind { Num, Int, Bool, Str, VecD, VecU };

struct Param {
    Kind kind;
    std::string value;    // textual value (default or parsed)
    bool set = false;
};

// trimLine of the reference: cut at '#' (comment) or keep a trailing '\'
// (continuation), strip white space
inline bool trim_line(std::string &s)
{
    size_t p = 0;
    while (p < s.size() && s[p] != '#' && s[p] != '\\') p++;
    const bool cont = p < s.size() && s[p] == '\\';
    s = s.substr(0, p);
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    s = (a == std::string::npos) ? "" : s.substr(a, b - a + 1);
    return cont;
}

inline bool read_line(FILE *f, std::string &out)
{
    char buf[8192];
    if (!std::fgets(buf, sizeof buf, f)) return false;
    std::string line = buf;
    bool cont = trim_line(line);
    while (cont && std::fgets(buf, sizeof buf, f)) {
        std::string next = buf;
        cont = trim_line(next);
        line += next.empty() ? "" : " " + next;
    }
    out = line;
    return true;
}

class Registry {
  public:
    // group -> name -> parameter
    std::map<std::string, std::map<std::string, Param>> g;
    // groups that accept any name (KspaceRegions): group -> name -> value, in name order
    std::map<std::string, std::map<std::string, std::string>> custom;
    std::vector<std::string> files;

    void add(const std::string &group, const std::string &name, Kind k, const std::string &def)
    {
        g[group][name] = Param{k, def, false};
    }
    void allow_custom(const std::string &group) { custom[group]; }

    void assign(const std::string &group, const std::string &name, const std::string &value, const std::string &where)
    {
        auto gi = g.find(group);
        if (gi == g.end()) throw ParamError{where + "Subgroup '" + group + "' not known"};
        auto pi = gi->second.find(name);
        if (pi == gi->second.end()) {
            auto ci = custom.find(group);
            if (ci != custom.end()) {
                ci->second[name] = value;
                return;
            }
            throw ParamError{where + "Name '" + name + "' not known"};
        }
        check(pi->second.kind, value, where + group + "." + name + ": ");
        pi->second.value = value;
        pi->second.set = true;
    }

    // one input file; the current group starts at the root (no group)
    void parse_file(const std::string &path)
    {
        FILE *f = std::fopen(path.c_str(), "r");
        if (!f) throw ParamError{"Could not open '" + path + "'"};
        files.push_back(path);
        parse_stream<void (*)(const std::string &)>(f, "", nullptr);
        std::fclose(f);
    }

    // lines of a file into `group`; section lines [name] call on_section
    // (multi-configuration files, code/src/io/ParameterIO.cpp:298-336)
    template <class F>
    void parse_stream(FILE *f, std::string group, F *on_section)
    {
        static const std::regex re_param(R"(((?:\w+\.)*)(\w+)\s*=\s*(.*))");
        static const std::regex re_sect(R"(\[(\.|:)?((?:\w+\.)*)(\w+)\])");
        std::string line;
        int ln = 0;
        while (read_line(f, line)) {
            ln++;
            std::smatch m;
            if (std::regex_match(line, m, re_param)) {
                std::string grp = m[1].str().empty() ? group : m[1].str().substr(0, m[1].str().size() - 1);
                assign(grp, m[2].str(), m[3].str(), "Line " + std::to_string(ln) + ": ");
            } else if (std::regex_match(line, m, re_sect)) {
                if (on_section) {
                    (*on_section)(line.substr(1, line.size() - 2));
                } else {
                    group = m[2].str() + m[3].str();
                    if (!g.count(group)) throw ParamError{"subgroup '" + group + "' not known"};
                }
            }
        }
    }
    void parse_file_into(const std::string &path, const std::string &group)
    {
        FILE *f = std::fopen(path.c_str(), "r");
        if (!f) throw ParamError{"Could not open '" + path + "'"};
        parse_stream<void (*)(const std::string &)>(f, group, nullptr);
        std::fclose(f);
    }

    // -Group.name=value
    void parse_option(const std::string &arg)
    {
        static const std::regex re(R"(-((?:\w+\.)*)(\w+)\s*=\s*(.*))");
        std::smatch m;
        if (!std::regex_match(arg, m, re)) throw ParamError{"Could not parse argument '" + arg + "'"};
        std::string grp = m[1].str().empty() ? "" : m[1].str().substr(0, m[1].str().size() - 1);
        assign(grp, m[2].str(), m[3].str(), "");
    }

    const Param &at(const std::string &group, const std::string &name) const { return g.at(group).at(name); }
    bool is_set(const std::string &group, const std::string &name) const { return at(group, name).set; }
    double num(const std::string &group, const std::string &name) const
    {
        double v = 0;
        std::sscanf(at(group, name).value.c_str(), "%lg", &v);
        return v;
    }
    long integer(const std::string &group, const std::string &name) const
    {
        long v = 0;
        std::sscanf(at(group, name).value.c_str(), "%ld", &v);
        return v;
    }
    bool flag(const std::string &group, const std::string &name) const { return parse_bool(at(group, name).value); }
    std::string str(const std::string &group, const std::string &name) const { return at(group, name).value; }
    std::vector<double> vec(const std::string &group, const std::string &name) const
    {
        std::vector<double> out;
        std::string s = at(group, name).value;
        char *copy = strdup(s.c_str());
        for (char *t = std::strtok(copy, ", "); t; t = std::strtok(nullptr, ", ")) {
            double v;
            if (std::sscanf(t, "%lg", &v) == 1) out.push_back(v);
        }
        std::free(copy);
        return out;
    }

    // current values of every parameter (to restore between multi-run configurations)
    std::map<std::string, std::map<std::string, Param>> snapshot() const { return g; }
    void restore(const std::map<std::string, std::map<std::string, Param>> &s) { g = s; }

    static bool parse_bool(const std::string &v)
    {
        if (v == "false") return false;
        if (v == "true") return true;
        int x = 0;
        std::sscanf(v.c_str(), "%d", &x);
        return x != 0;
    }

  private:
    static void check(Kind k, const std::string &v, const std::string &where)
    {
        double d;
        long l;
        switch (k) {
        case Kind::Num:
            if (std::sscanf(v.c_str(), "%lg", &d) != 1) throw ParamError{where + "Could not read value of '" + v + "'"};
            break;
        case Kind::Int:
            if (std::sscanf(v.c_str(), "%ld", &l) != 1) throw ParamError{where + "Could not read value of '" + v + "'"};
            break;
        case Kind::Bool:
            if (v != "true" && v != "false" && std::sscanf(v.c_str(), "%ld", &l) != 1)
                throw ParamError{where + "Could not interpret '" + v + "' as boolean"};
            break;
        case Kind::VecD:
        case Kind::VecU: {
            char *copy = strdup(v.c_str());
            for (char *t = std::strtok(copy, ", "); t; t = std::strtok(nullptr, ", "))
                if (std::sscanf(t, "%lg", &d) != 1) {
                    std::string bad = t;
                    std::free(copy);
                    throw ParamError{where + "Could not read value of '" + bad + "'"};
                }
            std::free(copy);
            break;
        }
        case Kind::Str:
            break;
        }
    }
};

// the parameters of the SBE run mode (names and defaults of
// code/src/io/{General,TightBinding,FourierTransform,Diagonalization,
// Propagation,Pulse,KspaceRegion,Output,Parallel}Parameter.cpp)
inline void register_sbe(Registry &r)
{
    using K = Kind;
    r.g[""];  // root
    r.add("General", "runMode", K::Str, "Dry");
    r.add("General", "dryRun", K::Bool, "false");
    r.add("General", "dimensionality", K::Int, "3");
    r.add("General", "verbose", K::Bool, "false");
    r.add("General", "veryVerbose", K::Bool, "false");
    r.add("General", "timeIt", K::Bool, "false");
    r.add("General", "finishMultiRuns", K::Bool, "false");
    r.add("Parallelization", "maxThreadCount", K::Int, "1000000");
    r.add("Parallelization", "rmaPollTime", K::Num, "0.1");
    r.add("Parallelization", "independentMpiIntegrations", K::Bool, "false");
    r.add("Parallelization", "useGPU", K::Bool, "false");
    r.add("TightBinding", "wannierSeed", K::Str, "");
    r.add("TightBinding", "symmetrizeHamiltonian", K::Bool, "true");
    r.add("TightBinding", "onlyRealMatrixElements", K::Bool, "false");
    r.add("TightBinding", "occupiedBands", K::Int, "0");
    r.add("TightBinding", "occupiedBelowBandGap", K::Bool, "false");
    r.add("TightBinding", "fermiLevel", K::Num, "0");
    r.add("TightBinding", "temperature", K::Num, "300");
    r.add("FourierTransform", "performDirectCalculation", K::Bool, "false");
    r.add("FourierTransform", "whTransformTestField", K::Bool, "false");
    r.add("FourierTransform", "saveIndices", K::VecU, "");
    r.add("FourierTransform", "fftPlanningMode", K::Str, "measure");
    r.add("FourierTransform", "planningTimeLimit", K::Num, "0");
    r.add("FourierTransform", "Nf", K::VecU, "1");
    r.add("FourierTransform", "maxNfPrime", K::Int, "5");
    r.add("Diagonalization", "mode", K::Str, "jacobi");
    r.add("Diagonalization", "maxSweepCount", K::Int, "5");
    r.add("Diagonalization", "maxRelError", K::Num, "1e-14");
    r.add("Propagation", "outputCount", K::Int, "500");
    r.add("Propagation", "Nk", K::VecU, "1, 1, 1");
    r.add("Propagation", "allowNkAdaption", K::Bool, "false");
    r.add("Propagation", "kGlobalShift", K::VecD, "0, 0, 0");
    r.add("Propagation", "epsAbs", K::Num, "1e-10");
    r.add("Propagation", "epsRel", K::Num, "1e-10");
    r.add("Propagation", "T2", K::Num, "-1");
    r.add("Propagation", "soothingWidth", K::Num, "0.025");
    r.add("Propagation", "occupationSmearingWidth", K::Num, "0.025");
    r.add("Pulse", "multiRunFile", K::Str, "");
    const char *pk[][3] = {{"pol", "v", ""}, {"type", "s", ""}, {"Emax", "n", "0"}, {"lambda", "n", "0"},
                           {"cep", "n", "0"}, {"FWHM", "n", "0"}, {"tCentral", "n", "0"}, {"tStart", "n", "0"},
                           {"tEnd", "n", "0"}, {"cyclesOn", "i", "3"}, {"risingCycles", "n", "1"}};
    auto kind_of = [](const char *c) {
        return c[0] == 'v' ? K::VecD : c[0] == 's' ? K::Str : c[0] == 'i' ? K::Int : K::Num;
    };
    // defaults of the default pulse: cep, tCentral, cyclesOn and risingCycles
    // have defaults (code/src/io/PulseParameter.cpp:38-62); the others must be
    // given (empty here, reported when a pulse needs them)
    const char *dflt[] = {"", "", "", "", "0", "", "0", "", "", "3", "1"};
    for (int j = 0; j < 11; j++) r.add("Pulse", pk[j][0], kind_of(pk[j][1]), dflt[j]);
    for (int i = 0; i < 10; i++)
        for (int j = 0; j < 11; j++) r.add("Pulse", pk[j][0] + std::to_string(i), kind_of(pk[j][1]), pk[j][2]);
    r.add("KspaceRegions", "storeDensityMatrix", K::Bool, "false");
    r.add("KspaceRegions", "densityOutputStride", K::Int, "1");
    r.allow_custom("KspaceRegions");
    r.add("Output", "disableRunSaving", K::Bool, "false");
    r.add("Output", "saveInput", K::Bool, "true");
    r.add("Output", "saveWannierInput", K::Bool, "false");
    r.add("Output", "outputDirectory", K::Str, "runs");
    r.add("Output", "saveAsNpz", K::Bool, "true");
    r.add("Output", "run", K::Str, "");
}

}  // namespace cli
