/*
 * phx_config.cpp -- PHOENIX flag parsing, envelopes, derived constants.
 * See phx_config.hpp for the reference units this restates.
 */
#include "phx_config.hpp"

#include <strings.h>

#include "phx_tableaux.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <random>
#include <sys/stat.h>

namespace phx {

namespace {

struct Args {
    std::vector<std::string> v;
    bool fp32 = false;
    std::string err;
    int argc() const { return (int)v.size(); }
    /* CLIO::findInArgv (commandline_io.cpp:110-126) */
    int find(const std::string& exact, int start = 0) const {
        for (int i = start; i < argc(); i++)
            if (v[i] == exact) return i;
        return -1;
    }
    int find(const std::vector<std::string>& keys, int start, const std::string& prefix) const {
        for (int i = start; i < argc(); i++)
            for (auto& k : keys)
                if (v[i] == prefix + k) return i;
        return -1;
    }
    /* CLIO::getNextInput (commandline_io.cpp:129-144): stod, rounded to the build's real */
    double num(const std::string& name, int& index) {
        if (index >= argc()) return 0.0;
        double r = 0.0;
        try {
            size_t pos = 0;
            r = std::stod(v[index], &pos);
        } catch (...) {
            if (err.empty()) err = "Invalid argument for '" + name + "': '" + v[index] + "' cannot be converted to a numerical value";
        }
        index++;
        if (fp32) r = (double)(float)r;
        return r;
    }
    /* CLIO::getNextStringInput (commandline_io.cpp:146-158) */
    std::string str(int& index) {
        if (index >= argc()) return "";
        return v[index++];
    }
};

template <class E>
bool cast_list(const std::string& input, const std::map<std::string, E>& table, uint32_t& out) {
    out = 0;
    size_t prev = 0, pos;
    while (true) {
        pos = input.find('+', prev);
        std::string sub = input.substr(prev, pos == std::string::npos ? std::string::npos : pos - prev);
        auto it = table.find(sub);
        if (it == table.end()) return false;
        out |= (uint32_t)it->second;
        if (pos == std::string::npos) break;
        prev = pos + 1;
    }
    return true;
}

const std::map<std::string, uint32_t> kBehavior = {{"add", BEH_ADD}, {"multiply", BEH_MULTIPLY}, {"replace", BEH_REPLACE}, {"adaptive", BEH_ADAPTIVE}, {"complex", BEH_COMPLEX}};
const std::map<std::string, uint32_t> kPol = {{"plus", POL_PLUS}, {"minus", POL_MINUS}, {"both", POL_BOTH}};
const std::map<std::string, uint32_t> kType = {{"gauss", ENV_GAUSS}, {"outerExponent", ENV_OUTER_EXPONENT}, {"ring", ENV_RING}, {"noDivide", ENV_NO_DIVIDE}, {"local", ENV_LOCAL}};
const std::map<std::string, uint32_t> kTemporal = {{"gauss", TMP_GAUSS}, {"iexp", TMP_IEXP}, {"osc", TMP_IEXP}, {"cos", TMP_COS}, {"constant", TMP_CONSTANT}};

std::string to_str6(double x) {
    char b[64];
    std::snprintf(b, sizeof b, "%f", x); /* std::to_string formatting */
    return b;
}

/* Envelope::addTemporal (envelope.cpp:89-120): groups keyed by their parameters */
std::string add_temporal(Envelope& e, double t0, double sigma, double freq, const std::string& s_temp) {
    std::string id = to_str6(t0) + to_str6(sigma) + to_str6(freq) + s_temp;
    if (!e.group_ids.count(id)) {
        Envelope::Temporal T;
        if (!cast_list(s_temp, kTemporal, T.kind)) return "unknown temporal envelope type '" + s_temp + "'";
        T.t0 = t0;
        T.sigma = sigma;
        T.freq = freq;
        int g = (int)e.group_ids.size();
        e.group_ids[id] = g;
        e.groups.push_back(T);
    }
    e.comps.back().group = e.group_ids[id];
    return "";
}

std::string add_temporal_loaded(Envelope& e, const std::string& path) {
    if (!e.group_ids.count(path)) {
        Envelope::Temporal T;
        T.kind = TMP_LOADED;
        T.load_path = path;
        /* time points: three rows t / re / im in a whitespace separated text file */
        std::ifstream f(path);
        if (f) {
            std::string line;
            std::vector<std::vector<double>> rows;
            while (std::getline(f, line)) {
                if (line.empty() || line[0] == '#') continue;
                std::istringstream is(line);
                std::vector<double> r;
                double x;
                while (is >> x) r.push_back(x);
                if (!r.empty()) rows.push_back(r);
            }
            if (rows.size() >= 2) {
                T.tp = rows[0];
                T.vr = rows[1];
                T.vi = rows.size() > 2 ? rows[2] : std::vector<double>(T.tp.size(), 0.0);
            }
        } else {
            std::fprintf(stderr, "[phx] Unable to load '%s'\n", path.c_str());
        }
        int g = (int)e.group_ids.size();
        e.group_ids[path] = g;
        e.groups.push_back(T);
    }
    e.comps.back().group = e.group_ids[path];
    return "";
}

/* Envelope::fromCommandlineArguments (envelope.cpp:148-315) */
std::string parse_envelope(Args& A, const std::vector<std::string>& keys, bool time, Envelope& ret) {
    int index = 0;
    const std::string key = keys.front();
    while ((index = A.find(keys, index, "--")) != -1) {
        ++index;
        std::string first = A.str(index);
        bool temporal_done = false;
        if (first == "load") {
            Envelope::Spatial s;
            s.load_path = A.str(index);
            s.amp = A.num(key + "_amp", index);
            std::string beh = A.str(index), pol = A.str(index);
            if (!cast_list(beh, kBehavior, s.behavior)) return "unknown behaviour '" + beh + "' for --" + key;
            if (!cast_list(pol, kPol, s.pol)) return "unknown polarization '" + pol + "' for --" + key;
            s.type = ENV_GAUSS;
            ret.comps.push_back(s);
        } else {
            if (!first.empty()) index--;
            Envelope::Spatial s;
            s.amp = A.num(key + "_amp", index);
            std::string sbeh = A.str(index);
            std::string peek = A.str(index);
            const bool named = peek == "width" || peek == "pos" || peek == "pol" || peek == "exponent" || peek == "charge" ||
                               peek == "type" || peek == "momenta" || peek == "time";
            std::string spol, sm, stype;
            if (named) {
                spol = "both";
                s.exponent = 1.0;
                sm = "none";
                stype = "gauss+noDivide";
                bool th = false;
                std::string t_type = "constant", t_load;
                double t0 = 0, sig = 0, fr = 0;
                auto kw = [&](const std::string& k) -> bool {
                    if (k == "width") {
                        s.width_x = A.num(key + "_width_x", index);
                        s.width_y = A.num(key + "_width_y", index);
                    } else if (k == "pos") {
                        s.x = A.num(key + "_X", index);
                        s.y = A.num(key + "_Y", index);
                    } else if (k == "pol") {
                        spol = A.str(index);
                    } else if (k == "exponent") {
                        s.exponent = A.num(key + "_exponent", index);
                    } else if (k == "charge") {
                        sm = A.str(index);
                    } else if (k == "type") {
                        stype = A.str(index);
                    } else if (k == "momenta") {
                        s.k0_x = A.num(key + "_k0_x", index);
                        s.k0_y = A.num(key + "_k0_y", index);
                    } else if (k == "time") {
                        th = true;
                        std::string tp = A.str(index);
                        if (tp == "load") {
                            t_load = A.str(index);
                        } else {
                            t_type = tp;
                            t0 = A.num(key + "_t0", index);
                            sig = A.num(key + "_sigma", index);
                            fr = A.num(key + "_freq", index);
                        }
                    } else
                        return false;
                    return true;
                };
                kw(peek);
                while (true) {
                    int before = index;
                    std::string k = A.str(index);
                    if (!kw(k)) {
                        if (index != before) index--;
                        break;
                    }
                }
                if (!cast_list(stype, kType, s.type)) return "unknown envelope type '" + stype + "' for --" + key;
                if (!cast_list(spol, kPol, s.pol)) return "unknown polarization '" + spol + "' for --" + key;
                if (!cast_list(sbeh, kBehavior, s.behavior)) return "unknown behaviour '" + sbeh + "' for --" + key;
                s.m = (sm == "none" || sm == "None") ? 0 : std::atoi(sm.c_str());
                ret.comps.push_back(s);
                std::string e2;
                if (th && !t_load.empty())
                    e2 = add_temporal_loaded(ret, t_load);
                else if (th)
                    e2 = add_temporal(ret, t0, sig, fr, t_type);
                else
                    e2 = add_temporal(ret, 0, 0, 0, "constant");
                if (!e2.empty()) return e2;
                temporal_done = true;
            } else {
                if (!peek.empty()) index--;
                s.width_x = A.num(key + "_width_x", index);
                s.width_y = A.num(key + "_width_y", index);
                s.x = A.num(key + "_X", index);
                s.y = A.num(key + "_Y", index);
                spol = A.str(index);
                s.exponent = A.num(key + "_exponent", index);
                sm = A.str(index);
                stype = A.str(index);
                int before = index;
                std::string mp = A.str(index);
                if (mp == "momenta") {
                    s.k0_x = A.num(key + "_k0_x", index);
                    s.k0_y = A.num(key + "_k0_y", index);
                } else if (index != before) {
                    index--;
                }
                if (!cast_list(stype, kType, s.type)) return "unknown envelope type '" + stype + "' for --" + key;
                if (!cast_list(spol, kPol, s.pol)) return "unknown polarization '" + spol + "' for --" + key;
                if (!cast_list(sbeh, kBehavior, s.behavior)) return "unknown behaviour '" + sbeh + "' for --" + key;
                s.m = (sm == "none" || sm == "None") ? 0 : std::atoi(sm.c_str());
                ret.comps.push_back(s);
            }
        }
        if (temporal_done) continue;
        int before = index;
        std::string next = A.str(index);
        if (!time || next != "time") {
            std::string e2 = add_temporal(ret, 0, 0, 0, "constant");
            if (!e2.empty()) return e2;
            if (index != before) index--;
            continue;
        }
        int b2 = index;
        if (A.str(index) == "load") {
            std::string e2 = add_temporal_loaded(ret, A.str(index));
            if (!e2.empty()) return e2;
        } else {
            if (index != b2) index--;
            std::string st = A.str(index);
            double t0 = A.num(key + "_t0", index);
            double sg = A.num(key + "_sigma", index);
            double fr = A.num(key + "_freq", index);
            std::string e2 = add_temporal(ret, t0, sg, fr, st);
            if (!e2.empty()) return e2;
        }
    }
    return A.err;
}

bool file_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

/* readConfigFromFile (config_readin.cpp:7-42) */
std::vector<std::string> expand_config(const std::vector<std::string>& in) {
    std::vector<std::string> buf = in;
    int index = -1;
    for (int i = 0; i < (int)in.size(); i++)
        if (in[i] == "--config") {
            index = i;
            break;
        }
    if (index < 0) return buf;
    while (true) {
        ++index;
        if (index >= (int)in.size()) break;
        std::string cfg = in[index++];
        if (cfg.empty()) break;
        if (!file_exists(cfg)) {
            std::fprintf(stderr, "[phx] Config file '%s' does not exist. Skipping.\n", cfg.c_str());
            return buf;
        }
        std::ifstream f(cfg);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream is(line);
            std::string w;
            while (is >> w) {
                if (w[0] == '#') break;
                buf.push_back(w);
            }
            buf.push_back(" ");
        }
    }
    return buf;
}

} // namespace

bool Envelope::allConstant() const {
    for (auto& g : groups)
        if (!(g.kind & TMP_CONSTANT)) return false;
    return true;
}

/* Envelope::updateTemporal (envelope.cpp:466-499) with the oscillators of envelope.hpp:165-175 */
std::complex<double> Envelope::temporal(int g, double t) const {
    const Temporal& T = groups[g];
    std::complex<double> v = 1.0;
    if (T.kind & TMP_CONSTANT) return v;
    if (T.kind & TMP_LOADED) {
        if (T.tp.empty()) return v;
        if (t > T.tp.back()) return {T.vr.back(), T.vi.back()};
        auto it = std::lower_bound(T.tp.begin(), T.tp.end(), t);
        size_t idx = std::min<size_t>(std::max<size_t>(1, (size_t)(it - T.tp.begin())), T.tp.size() - 1);
        const double t1 = T.tp[idx - 1], t2 = T.tp[idx];
        std::complex<double> v1(T.vr[idx - 1], T.vi[idx - 1]), v2(T.vr[idx], T.vi[idx]);
        v = v1 + (v2 - v1) * (t - t1) / (t2 - t1);
    }
    if (T.kind & TMP_IEXP) {
        v = std::exp(-std::complex<double>((t - T.t0) * (t - T.t0) / (2.0 * T.sigma * T.sigma), T.freq * (t - T.t0)));
    } else if (T.kind & TMP_COS) {
        const double p = (t - T.t0) / T.sigma;
        v = std::exp(-0.5 * p * p) * (1.0 + std::cos(T.freq * (t - T.t0))) / 2.0;
    } else if (T.kind & TMP_GAUSS) {
        const double p = (t - T.t0) / T.sigma;
        v = std::exp(-0.5 * std::pow(p * p, T.freq));
    }
    return v;
}

bool load_txt_matrix(const std::string& path, std::vector<std::complex<double>>& out, int& N_c, int& N_r) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    std::string line, tok;
    std::getline(f, line);
    std::istringstream hs(line);
    unsigned nc = 0, nr = 0;
    hs >> tok >> tok >> nc >> nr;
    N_c = (int)nc;
    N_r = (int)nr;
    const size_t N = (size_t)nc * nr;
    out.assign(N, {0.0, 0.0});
    size_t i = 0;
    while (std::getline(f, line)) {
        if (line.size() < 1 || line[0] == '#') continue;
        std::istringstream is(line);
        double x;
        while (is >> x) {
            if (i < N)
                out[i] = {x, 0.0};
            else if (i - N < N)
                out[i - N] = {out[i - N].real(), x};
            i++;
        }
    }
    return true;
}

std::string Envelope::calculate(std::complex<double>* buffer, int group, uint32_t polarization, int N_c, int N_r, double L_x,
                                double L_y, double dx, double dy, double default_if_unset) const {
    /* file-backed components are read first (envelope.hpp:121-135); a missing file leaves zeros */
    std::vector<std::vector<std::complex<double>>> cache(comps.size());
    for (size_t c = 0; c < comps.size(); c++) {
        if (comps[c].load_path.empty()) continue;
        int nc = 0, nr = 0;
        if (!load_txt_matrix(comps[c].load_path, cache[c], nc, nr)) {
            std::fprintf(stderr, "[phx] Unable to load '%s'\n", comps[c].load_path.c_str());
            cache[c].assign((size_t)N_c * N_r, {0.0, 0.0});
        } else if ((size_t)nc * nr != (size_t)N_c * N_r) {
            cache[c].resize((size_t)N_c * N_r, {0.0, 0.0});
        }
    }
    for (int row = 0; row < N_r; row++)
        for (int col = 0; col < N_c; col++) {
            const size_t i = (size_t)row * N_c + col;
            buffer[i] = {0.0, 0.0};
            bool set = false;
            for (size_t c = 0; c < comps.size(); c++) {
                const Spatial& s = comps[c];
                if (group >= 0 && s.group != group) continue;
                if (s.pol != POL_BOTH && s.pol != polarization && polarization != POL_BOTH) continue;
                double cx = -L_x / 2.0 + dx * col;
                double cy = -L_y / 2.0 + dy * row;
                if (s.type & ENV_LOCAL) {
                    cx = -1.0 + 2.0 * col / (N_c - 1);
                    cy = -1.0 + 2.0 * row / (N_r - 1);
                }
                set = true;
                std::complex<double> amplitude(s.amp, 0.0);
                double pre = 1, ef = 1;
                std::complex<double> charge(1.0, 0.0);
                if (!s.load_path.empty()) {
                    amplitude = cache[c][i] * s.amp;
                } else {
                    const double ax = (cx - s.x) / s.width_x, ay = (cy - s.y) / s.width_y;
                    const double ex = 0.5 * (ax * ax + ay * ay);
                    ef = std::exp(-std::pow(ex, s.exponent));
                    if (s.type & ENV_OUTER_EXPONENT) ef = std::pow(std::exp(-ex), s.exponent);
                    pre = (s.type & ENV_RING) ? ex : 1.0;
                    charge = std::exp(std::complex<double>(0.0, s.m * std::atan2(cx - s.x, cy - s.y)));
                    if (!(s.type & ENV_NO_DIVIDE)) amplitude = amplitude / std::sqrt(2 * 3.1415 * s.width_x * s.width_y);
                }
                const std::complex<double> mom = std::exp(std::complex<double>(0.0, s.k0_x * (cx - s.x) + s.k0_y * (cy - s.y)));
                if (s.behavior & BEH_ADAPTIVE) amplitude = s.amp * buffer[i];
                if (s.behavior & BEH_COMPLEX) amplitude = std::complex<double>(0.0, amplitude.real());
                const std::complex<double> contribution = amplitude * pre * ef * charge * mom;
                if (s.behavior & BEH_ADD)
                    buffer[i] = buffer[i] + contribution;
                else if (s.behavior == BEH_MULTIPLY)
                    buffer[i] = buffer[i] * contribution;
                else if (s.behavior == BEH_REPLACE)
                    buffer[i] = contribution;
            }
            if (!set) buffer[i] = {default_if_unset, 0.0};
        }
    return "";
}

std::string parse_config(SystemConfig& c, const std::vector<std::string>& argv_in, bool fp32) {
    Args A;
    A.v = expand_config(argv_in);
    A.fp32 = fp32;
    c.fp32 = fp32;
    int index = 0;
    /* system_initialization.cpp:13-261, in the reference's order */
    c.tetm = A.find("-tetm") != -1;
    if ((index = A.find({"gammaC", "gamma_C", "gammac", "gamma_c"}, 0, "--")) != -1) c.gamma_c = A.num("gamma_c", ++index);
    if ((index = A.find({"gammaR", "gamma_R", "gammar", "gamma_r"}, 0, "--")) != -1) c.gamma_r = A.num("gamma_r", ++index);
    if ((index = A.find({"gc", "g_c"}, 0, "--")) != -1) c.g_c = A.num("g_c", ++index);
    if ((index = A.find({"gr", "g_r"}, 0, "--")) != -1) c.g_r = A.num("g_r", ++index);
    if ((index = A.find("--R")) != -1) c.R = A.num("R", ++index);
    if ((index = A.find({"L", "gridlength", "xmax"}, 0, "--")) != -1) {
        c.L_x = A.num("L", ++index);
        c.L_y = A.num("L", index);
    }
    if ((index = A.find({"g_pm", "gpm", "g_PM"}, 0, "--")) != -1) c.g_pm = A.num("gm", ++index);
    if ((index = A.find({"deltaLT", "delta_LT", "deltalt", "dlt"}, 0, "--")) != -1) c.delta_LT = A.num("deltaLT", ++index);
    if ((index = A.find("--threads")) != -1) c.threads = (int)A.num("threads", ++index);
    if ((index = A.find("--blocksize")) != -1) c.block_size = (int)A.num("block_size", ++index);
    if ((index = A.find({"output", "fileoutput", "FO"}, 0, "--")) != -1) {
        c.output_keys.clear();
        std::string s = A.str(++index);
        size_t prev = 0, pos;
        while ((pos = s.find(',', prev)) != std::string::npos) {
            c.output_keys.push_back(s.substr(prev, pos - prev));
            prev = pos + 1;
        }
        c.output_keys.push_back(s.substr(prev));
    }
    if ((index = A.find({"of", "outputFileType"}, 0, "--")) != -1) {
        std::string t = A.str(++index);
        if (t == "phoenix") t = "ash";
        if (t == "numpy") t = "npy";
        if (t == "text") t = "txt";
        if (t == "matlab") t = "mat";
        if (t != "ash" && t != "npy" && t != "txt" && t != "mat") {
            std::fprintf(stderr, "[phx] Invalid output file type '%s'. Defaulting to 'txt'...\n", t.c_str());
            t = "txt";
        }
        c.output_filetype = t;
    }
    if ((index = A.find({"fileout-precision", "fileoutPrecision"}, 0, "--")) != -1)
        c.fileout_precision = (int)A.num("fileout_precision", ++index);
    if ((index = A.find({"historyMatrix", "hM"}, 0, "--")) != -1) {
        c.history_x0 = (int)A.num("history_matrix_start_x", ++index);
        c.history_x1 = (int)A.num("history_matrix_end_x", index);
        c.history_y0 = (int)A.num("history_matrix_start_y", index);
        c.history_y1 = (int)A.num("history_matrix_end_y", index);
        c.history_increment = (int)A.num("history_matrix_output_increment", index);
        c.history = true;
    }
    if ((index = A.find({"historyTime", "hT"}, 0, "--")) != -1) {
        c.history_start = A.num("history_time", ++index);
        c.history_every = (int)A.num("history_time_every", index);
    }
    if ((index = A.find({"N", "gridsize"}, 0, "--")) != -1) {
        c.N_c = (int)A.num("N_c", ++index);
        c.N_r = (int)A.num("N_r", index);
    }
    c.subgrids_columns = c.subgrids_rows = 0;
    if ((index = A.find({"subgrids", "sg"}, 0, "--")) != -1) {
        c.subgrids_columns = (int)A.num("subgrids_columns", ++index);
        c.subgrids_rows = (int)A.num("subgrids_rows", index);
    }
    if ((index = A.find({"tmax", "tend"}, 0, "--")) != -1) c.t_max = A.num("s_t_max", ++index);
    if ((index = A.find({"tstep", "dt"}, 0, "--")) != -1) {
        c.dt = A.num("t_step", ++index);
        c.do_overwrite_dt = false;
    }
    if ((index = A.find("--tol")) != -1) c.tolerance = A.num("tol", ++index);
    if ((index = A.find("--rkvdt")) != -1) {
        c.dt_min = A.num("dt_min", ++index);
        c.dt_max = A.num("dt_max", index);
    }
    c.imag_time_amplitude = 0.0;
    if ((index = A.find("--imagTime")) != -1) c.imag_time_amplitude = A.num("imag_time_amplitude", ++index);
    if ((index = A.find("--fftEvery")) != -1) c.fft_every = A.num("fft_every", ++index);
    c.iterator = "RK4";
    if (A.find("-ssfm") != -1) c.iterator = "SSFM";
    if ((index = A.find("--iterator")) != -1) c.iterator = A.str(++index);
    c.adaptive = A.find({"adaptive", "adaptiveTimestep"}, 0, "-") != -1;
    if ((index = A.find({"initRandom", "iR"}, 0, "--")) != -1) {
        c.random_init = true;
        c.random_amp = A.num("random_system_amplitude", ++index);
        c.random_seed = 0;
        std::string sseed = A.str(index);
        if (sseed != "random") {
            try {
                c.random_seed = (uint32_t)std::stod(sseed);
            } catch (...) {
                return "Invalid argument for 'random_seed': '" + sseed + "'";
            }
        } else {
            c.random_seed = std::random_device{}();
        }
    }
    if ((index = A.find("--dw")) != -1) c.stochastic_amplitude = A.num("dw", ++index);
    if ((index = A.find("--outEvery")) != -1) c.output_every = A.num("output_every", ++index);
    c.periodic_x = c.periodic_y = false;
    if ((index = A.find("--boundary")) != -1) {
        std::string bx = A.str(++index), by = A.str(index);
        c.periodic_x = bx == "periodic";
        c.periodic_y = by == "periodic";
    }
    c.t = 0.0;
    if ((index = A.find("--t0")) != -1) c.t = A.num("t0", ++index);
    if ((index = A.find("--hbar")) != -1) c.h_bar = A.num("hbar", ++index);
    if ((index = A.find({"electron_charge", "e"}, 0, "--")) != -1) c.e_e = A.num("e", ++index);
    if ((index = A.find({"me", "m_e", "electron_mass"}, 0, "--")) != -1) c.m_e = A.num("me", ++index);
    if ((index = A.find({"hbarscaled", "hbar_scaled", "hbars"}, 0, "--")) != -1) c.h_bar_s = A.num("hbars", ++index);
    if ((index = A.find({"meff", "m_eff"}, 0, "--")) != -1) c.m_eff = A.num("m_eff", ++index);
    /* output directory, file-name prefix, per-file log (filehandler.cpp:32-74) */
    if ((index = A.find({"path", "outputPath"}, 0, "--")) != -1) {
        c.path = A.str(++index);
        if (!c.path.empty() && c.path.back() != '/') c.path += "/";
    }
    if ((index = A.find("--name")) != -1) c.output_name = A.str(++index);
    if (A.find("-verboseIO") != -1) c.verbose_io = true;
#define PHX_X(name, def) \
    if ((index = A.find("--" #name)) != -1) c.name = A.num(#name, ++index);
    PHX_CUSTOM_PARAMETERS(PHX_X)
#undef PHX_X
    if (!A.err.empty()) return A.err;

    std::string e;
    if (!(e = parse_envelope(A, {"pump"}, true, c.pump)).empty()) return e;
    if (!(e = parse_envelope(A, {"potential"}, true, c.potential)).empty()) return e;
    if (!(e = parse_envelope(A, {"pulse"}, true, c.pulse)).empty()) return e;
    if (!(e = parse_envelope(A, {"fftMask"}, false, c.fft_mask)).empty()) return e;
    if (!(e = parse_envelope(A, {"initState", "initialState", "iS"}, false, c.initial_state)).empty()) return e;
    if (!(e = parse_envelope(A, {"initReservoir", "initialReservoir", "iR"}, false, c.initial_reservoir)).empty()) return e;

    c.use_reservoir = true;
    c.use_pumps = c.pump.size() > 0;
    c.use_pulses = c.pulse.size() > 0;
    c.use_potentials = c.potential.size() > 0;
    c.use_fft_mask = c.fft_mask.size() > 0;
    c.use_stochastic = c.stochastic_amplitude > 0.0;
    if (c.pump.size() == 0 && c.initial_reservoir.size() == 0) c.use_reservoir = false;
    if (A.find("-noReservoir") != -1) {
        c.use_reservoir = false;
        c.no_reservoir_flag = true;
    }
    return A.err;
}

std::string finalize_config(SystemConfig& c) {
    /* iterator registry (solver_factory.hpp:53-60 throws on unknown keys); keys match
       without regard to case.  An adaptive request on an iterator registered without
       variable time stepping runs it fixed ("Cannot use variable time step with this
       solver.", solver.cu:23-26). */
    int n_it = 0;
    const phx::IteratorInfo* table = phx::iterator_table(&n_it);
    const phx::IteratorInfo* found = nullptr;
    std::string available;
    for (int k = 0; k < n_it; k++) {
        if (strcasecmp(table[k].key, c.iterator.c_str()) == 0) found = &table[k];
        available += (k ? ", " : "") + std::string(table[k].key);
    }
    if (strcasecmp(c.iterator.c_str(), "SSFM") == 0) {
        /* split-step Fourier (split_step_fourier_method.cu:13-20): no halo, fixed step */
        c.iterator = "SSFM";
        c.adaptive = false;
        c.halo = 0;
    } else {
        if (!found)
            return "iterator '" + c.iterator + "' is not provided by this library (available: " + available + ", SSFM)";
        c.iterator = found->key;
        if (c.adaptive && !found->can_adapt) c.adaptive = false;
        c.halo = phx::rk_tableau_stages(c.adaptive ? found->adaptive : found->fixed);
    }
    /* imaginary time turns the reference's (complex-stored) reservoir complex: n advances by -i dt
       (runge_sum_to_input_k with complex_dt, cuda_macro.cuh:160-178).  The split-step step takes no
       -i dt (split_step_fourier_method.cu:25-71): there n stays real and is only normalised. */
    c.complex_reservoir = c.imag_time_amplitude != 0.0 && c.use_reservoir && c.iterator != "SSFM";
    /* calculateAuto (system.cpp:116-160) */
    if (c.h_bar_s < 0) c.h_bar_s = c.h_bar / c.e_e * 1E12;
    if (c.m_eff < 0) c.m_eff = c.m_e / c.e_e * 1E8;
    const double dt_scaling_factor = c.m_e / c.e_e * 1E8 / c.m_eff;
    c.dx = c.L_x / (c.N_c - 1);
    c.dy = c.L_y / (c.N_r - 1);
    c.dV = c.dx * c.dy;
    c.one_over_dx2 = 1.0 / (c.dx * c.dx);
    c.one_over_dy2 = 1.0 / (c.dy * c.dy);
    c.m2_over_dx2_p_dy2 = -2.0 * (c.one_over_dx2 + c.one_over_dy2);
    c.m_eff_scaled = -0.5 * c.h_bar_s * c.h_bar_s / c.m_eff;
    c.magic_timestep = 0.5 * c.dx * c.dy / dt_scaling_factor;
    if (c.do_overwrite_dt) c.dt = c.magic_timestep;
    c.one_over_h_bar_s = 1.0 / c.h_bar_s;
    if (c.fp32) {
        auto f = [](double& x) { x = (double)(float)x; };
        f(c.dx), f(c.dy), f(c.dV), f(c.one_over_dx2), f(c.one_over_dy2), f(c.m2_over_dx2_p_dy2), f(c.m_eff_scaled),
            f(c.one_over_h_bar_s), f(c.dt);
    }
    /* validateInputs (system_validate_input.cpp:19-56) */
    std::string bad;
    if (c.N_c <= 0 || c.N_r <= 0) bad += "N cannot be negative; ";
    else if (c.N_c % 2 != 0 || c.N_r % 2 != 0) bad += "Input Dimensions have to be even; ";
    if (c.t_max < 0) bad += "t_max cannot be negative; ";
    if (c.dt <= 0) bad += "dt cannot be negative or zero; ";
    if (c.dt > c.t_max) bad += "dt cannot be larger than t_max; ";
    if (c.dt_max < 0) bad += "dt_max cannot be negative; ";
    if (c.dt_min < 0) bad += "dt_min cannot be negative; ";
    if (!bad.empty()) return "Invalid input! " + bad;
    /* a complex temporal pump amplitude makes n complex (k_n += P_g a_g(w), kernel_gp_compute.cuh:70-75) */
    if (c.use_pumps && c.use_reservoir)
        for (int g = 0; g < c.pump.groupSize(); g++)
            if (c.pump.groups[g].kind & TMP_IEXP) c.complex_reservoir = true;
    return "";
}

template <typename real>
PhxPhys<real> make_phys(const SystemConfig& c) {
    PhxPhys<real> p{};
    p.m_eff_scaled = (real)c.m_eff_scaled;
    p.one_over_h_bar_s = (real)c.one_over_h_bar_s;
    p.m2_over_dx2_p_dy2 = (real)c.m2_over_dx2_p_dy2;
    p.one_over_dx2 = (real)c.one_over_dx2;
    p.one_over_dy2 = (real)c.one_over_dy2;
    p.dx = (real)c.dx;
    p.dy = (real)c.dy;
    p.dV = (real)c.dV;
    p.gamma_c = (real)c.gamma_c;
    p.gamma_r = (real)c.gamma_r;
    p.g_c = (real)c.g_c;
    p.g_r = (real)c.g_r;
    p.R = (real)c.R;
    p.g_pm = (real)c.g_pm;
    p.delta_LT = (real)c.delta_LT;
#define PHX_X(name, def) p.name = (real)c.name;
    PHX_CUSTOM_PARAMETERS(PHX_X)
#undef PHX_X
    phx_fold<real>(p, c.m_eff_scaled, c.one_over_h_bar_s, c.m2_over_dx2_p_dy2, c.one_over_dx2, c.one_over_dy2, c.dx, c.dy);
    return p;
}

template PhxPhys<float> make_phys<float>(const SystemConfig&);
template PhxPhys<double> make_phys<double>(const SystemConfig&);

} // namespace phx
