/*
 * phx_main.cpp -- PHOENIX_gpu_{fp32,fp64}: command line front end over libphx.
 *
 * Mirrors main() (source/main.cu:46-135): parse flags, initialise, write the initial
 * matrices (solver_output_matrices.cu:72-150), run the main loop one output interval at a
 * time -- caching t, min/max |psi| and the temporal amplitudes after each interval
 * (solver_cache.cu:7-40), writing --historyMatrix snapshots (solver_cache.cu:57-72) and
 * printing the progress block that pyphoenix.run relays (system_print_output.cpp:812-848) --
 * then write the final matrices and scalar.txt (solver_finalize.cu:5-11) in the reference's
 * txt format (filehandler.cpp:233-313) and print the runtime summary line that tooling
 * parses (system_print_output.cpp:731, "... mus/it").  What is written is selected by the
 * --output keys exactly as SystemParameters::doOutput does (system_parameters.hpp:117-119;
 * default "mat,scalar", system.cpp:65).  Errors exit with status 1
 * (system_validate_input.cpp:53-56).
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "phx.h"

#ifndef PHX_MAIN_PRECISION
#define PHX_MAIN_PRECISION 1
#endif

namespace {

struct Out {
    phx_output_config oc;
    std::string path;  /* --path, ending in '/' */
    std::string prefix; /* "<--name>_" or empty: every file is path + prefix + name (filehandler.cpp:461-463) */
    std::string ftype; /* txt, npy, ash, mat */
    bool fp32 = false; /* binary files hold the build's precision */
    phx_info in;
    std::vector<std::string> keys;
    std::map<std::string, std::vector<double>> scalars; /* cacheValues: ordered by key, like std::map there */
    int history_counter = 1;

    bool want(std::initializer_list<const char*> any) const {
        for (const char* k : any)
            if (std::find(keys.begin(), keys.end(), k) != keys.end()) return true;
        return false;
    }
    std::string num(double v) const {
        char b[64];
        if (oc.precision > 0)
            std::snprintf(b, sizeof b, "%.*g", oc.precision, v);
        else
            std::snprintf(b, sizeof b, "%.17g", v);
        return b;
    }
};

/* FileHandler::Header (filehandler.hpp:32-69) */
std::string header(const Out& o, int nc, int nr, double Lx, double Ly, double t, double t0 = 0, double freq = 0,
                   double sigma = 0) {
    std::ostringstream h;
    h << "# SIZE " << nc << " " << nr << " LX " << Lx << " LY " << Ly << " DX " << o.in.dx << " DY " << o.in.dy << " TIME "
      << t;
    if (t0 != 0 && freq != 0 && sigma != 0) h << " OSC T0 " << t0 << " FREQ " << freq << " SIGMA " << sigma;
    h << " :: PHOENIX_ MATRIX\n";
    return h.str();
}

/* Binary forms (simple_binary.hpp): npy ("\x93NUMPY" magic) and ash ("\x93PHOENIX" magic) share
   the NumPy header layout -- version 1.0, little-endian dtype ('<f4', '<f8', '<c8', '<c16'), C
   order, shape (rows, cols), the header text padded with spaces to (10 + length) % 16 == 15 and
   ended by a newline (simple_binary.hpp:51-84) -- followed by the raw data, complex values
   interleaved; mat is the reference's MATLAB-v4-style layout: a 128-byte text header, five int32
   (type 0, rows, cols, complex flag, name length), the name "mat", then the real block and (if
   complex) the imaginary block, column-major (simple_binary.hpp:145-262). */
template <typename real>
void write_binary(FILE* f, const std::string& ftype, const std::vector<real>& re, const std::vector<real>& im, bool cplx,
                  int rows, int cols) {
    const size_t n = (size_t)rows * cols;
    if (ftype == "mat") {
        char head[128] = {0};
        std::snprintf(head, sizeof head, "%s",
                      cplx ? "MATLAB 4.0 MAT-file, Created by PHOENIX::Output (complex)" : "MATLAB 4.0 MAT-file, Created by PHOENIX::Output");
        std::fwrite(head, 1, 128, f);
        const int32_t ah[5] = {0, rows, cols, cplx ? 1 : 0, 4};
        std::fwrite(ah, sizeof(int32_t), 5, f);
        std::fwrite("mat", 1, 4, f);
        std::vector<real> col(n);
        for (int part = 0; part < (cplx ? 2 : 1); part++) {
            const std::vector<real>& src = part == 0 ? re : im;
            for (int j = 0; j < cols; j++)
                for (int i = 0; i < rows; i++) col[(size_t)j * rows + i] = src[(size_t)i * cols + j];
            std::fwrite(col.data(), sizeof(real), n, f);
        }
        return;
    }
    std::string descr = std::string("<") + (cplx ? (sizeof(real) == 4 ? "c8" : "c16") : (sizeof(real) == 4 ? "f4" : "f8"));
    std::string h = "{'descr': '" + descr + "', 'fortran_order': False, 'shape': (" + std::to_string(rows) + ", " +
                    std::to_string(cols) + "), }";
    while ((10 + h.size()) % 16 != 15) h += ' ';
    h += '\n';
    const std::string magic = ftype == "npy" ? "NUMPY" : "PHOENIX";
    std::fputc(0x93, f);
    std::fwrite(magic.data(), 1, magic.size(), f);
    std::fputc(0x01, f);
    std::fputc(0x00, f);
    std::fputc((int)(h.size() & 0xFF), f);
    std::fputc((int)((h.size() >> 8) & 0xFF), f);
    std::fwrite(h.data(), 1, h.size(), f);
    if (!cplx) {
        std::fwrite(re.data(), sizeof(real), n, f);
        return;
    }
    std::vector<real> il(2 * n);
    for (size_t i = 0; i < n; i++) {
        il[2 * i] = re[i];
        il[2 * i + 1] = im[i];
    }
    std::fwrite(il.data(), sizeof(real), il.size(), f);
}

/* -verboseIO: one line per written matrix (filehandler.cpp:309-312) */
void log_write(const Out& o, const std::string& file, size_t elements) {
    if (o.oc.verbose_io) std::printf("Output %zu elements to '%s'.\n", elements, file.c_str());
}

/* outputMatrixToFile (filehandler.cpp:233-398): the region [x0, x1) x [y0, y1), every inc-th cell;
   txt: header, then the real parts and (complex matrices) the imaginary parts (filehandler.cpp:282-307) */
void write_matrix(const Out& o, const std::string& name, const std::vector<double>& z, int x0, int x1, int y0, int y1,
                  int inc, const std::string& head, bool cplx = true) {
    const std::string file = o.path + o.prefix + name + "." + o.ftype;
    FILE* f = std::fopen(file.c_str(), o.ftype == "txt" ? "w" : "wb");
    if (!f) {
        std::fprintf(stderr, "Unable to write '%s'\n", file.c_str());
        return;
    }
    const int N_c = o.in.N_c;
    if (o.ftype != "txt") {
        const int cols = (x1 - x0 + inc - 1) / inc, rows = (y1 - y0 + inc - 1) / inc;
        std::vector<double> re, im;
        for (int r = y0; r < y1; r += inc)
            for (int c = x0; c < x1; c += inc) {
                re.push_back(z[2 * ((size_t)r * N_c + c)]);
                im.push_back(z[2 * ((size_t)r * N_c + c) + 1]);
            }
        if (o.fp32)
            write_binary<float>(f, o.ftype, std::vector<float>(re.begin(), re.end()), std::vector<float>(im.begin(), im.end()),
                                cplx, rows, cols);
        else
            write_binary<double>(f, o.ftype, re, im, cplx, rows, cols);
        std::fclose(f);
        log_write(o, file, (size_t)rows * cols);
        return;
    }
    std::fputs(head.c_str(), f);
    for (int part = 0; part < (cplx ? 2 : 1); part++) /* real matrices: the real block only (filehandler.cpp:375-391) */
        for (int r = y0; r < y1; r += inc) {
            std::string line;
            for (int c = x0; c < x1; c += inc) line += o.num(z[2 * ((size_t)r * N_c + c) + part]) + " ";
            line += "\n";
            std::fputs(line.c_str(), f);
        }
    std::fclose(f);
    log_write(o, file, (size_t)((y1 - y0 + inc - 1) / inc) * ((x1 - x0 + inc - 1) / inc));
}

void write_full(const Out& o, const std::string& name, const std::vector<double>& z, bool cplx, double t, double t0 = 0,
                double freq = 0, double sigma = 0) {
    write_matrix(o, name, z, 0, o.in.N_c, 0, o.in.N_r, 1,
                 header(o, o.in.N_c, o.in.N_r, o.oc.L_x, o.oc.L_y, t, t0, freq, sigma), cplx);
}

/* a device field as complex doubles (real fields get a zero imaginary part) */
template <typename real>
bool field(phx_solver* s, int id, bool cplx, std::vector<double>& z) {
    phx_info in;
    phx_get_info(s, &in);
    const size_t cells = (size_t)in.N_c * in.N_r;
    std::vector<real> buf((cplx ? 2 : 1) * cells);
    if (phx_get_field(s, id, buf.data(), buf.size() * sizeof(real)) != PHX_OK) return false;
    z.assign(2 * cells, 0.0);
    for (size_t i = 0; i < cells; i++) {
        z[2 * i] = cplx ? (double)buf[2 * i] : (double)buf[i];
        z[2 * i + 1] = cplx ? (double)buf[2 * i + 1] : 0.0;
    }
    return true;
}

/* Solver::outputInitialMatrices (solver_output_matrices.cu:72-150) */
template <typename real>
void output_initial(phx_solver* s, const Out& o) {
    const char* pm[2] = {"plus", "minus"};
    const int ncomp = o.in.tetm ? 2 : 1;
    std::vector<double> z((size_t)2 * o.in.N_c * o.in.N_r);
    for (int k = 0; k < ncomp; k++) {
        const std::string sfx = pm[k];
        if (o.want({"all", "mat", k == 0 ? "initial_plus" : "initial_minus", "initial"})) {
            if (field<real>(s, k == 0 ? PHX_PSI_PLUS : PHX_PSI_MINUS, true, z)) write_full(o, "initial_wavefunction_" + sfx, z, true, o.in.t);
            if (o.in.use_reservoir && field<real>(s, k == 0 ? PHX_N_PLUS : PHX_N_MINUS, o.in.complex_reservoir != 0, z))
                write_full(o, "initial_reservoir_" + sfx, z, true, o.in.t);
        }
        struct Src {
            int id;
            const char* name;
            bool on;
        } srcs[3] = {{PHX_SRC_PUMP, "pump", o.in.use_reservoir && o.want({"all", "mat", k == 0 ? "pump_plus" : "pump_minus", "pump"})},
                     {PHX_SRC_PULSE, "pulse", o.want({"all", "mat", k == 0 ? "pulse_plus" : "pulse_minus", "pulse"})},
                     {PHX_SRC_POTENTIAL, "potential",
                      o.want({"all", "mat", k == 0 ? "potential_plus" : "potential_minus", "potential"})}};
        for (const auto& src : srcs) {
            int ng = 0;
            if (!src.on || phx_source_groups(s, src.id, &ng) != PHX_OK) continue;
            for (int g = 0; g < ng; g++) {
                double re, im, t0, fr, sg;
                int cst;
                phx_source_temporal(s, src.id, g, o.in.t, &re, &im, &t0, &fr, &sg, &cst);
                if (phx_source_field(s, src.id, k, g, z.data(), z.size() * sizeof(double)) != PHX_OK) continue;
                const std::string name = std::string(src.name) + "_" + sfx + (g > 0 ? "_" + std::to_string(g) : "");
                write_full(o, name, z, src.id == PHX_SRC_PULSE, o.in.t, t0, fr, sg);
            }
        }
        if (o.in.fft_every < o.in.t_max && o.in.use_fft_mask && o.want({"all", "mat", k == 0 ? "fft_plus" : "fft_minus", "fft"}))
            if (field<real>(s, k == 0 ? PHX_FFT_MASK_PLUS : PHX_FFT_MASK_MINUS, false, z)) write_full(o, "fft_mask_" + sfx, z, false, o.in.t);
    }
}

/* Solver::cacheValues (solver_cache.cu:7-40): t, min/max |psi| over the grid, the temporal
   amplitudes of every non-constant group */
template <typename real>
void cache_values(phx_solver* s, Out& o) {
    phx_get_info(s, &o.in);
    o.scalars["t"].push_back(o.in.t);
    std::vector<double> z;
    for (int k = 0; k < (o.in.tetm ? 2 : 1); k++) {
        if (!field<real>(s, k == 0 ? PHX_PSI_PLUS : PHX_PSI_MINUS, true, z)) continue;
        double mn = INFINITY, mx = 0.0;
        for (size_t i = 0; i < z.size() / 2; i++) {
            const double a = z[2 * i] * z[2 * i] + z[2 * i + 1] * z[2 * i + 1];
            mn = std::min(mn, a);
            mx = std::max(mx, a);
        }
        o.scalars[k == 0 ? "min_plus" : "min_minus"].push_back(std::sqrt(mn));
        o.scalars[k == 0 ? "max_plus" : "max_minus"].push_back(std::sqrt(mx));
    }
    const struct {
        int id;
        const char* name;
        bool complex_value;
    } srcs[3] = {{PHX_SRC_PULSE, "pulse", true}, {PHX_SRC_PUMP, "pump", false}, {PHX_SRC_POTENTIAL, "potential", false}};
    for (const auto& src : srcs) {
        int ng = 0;
        if (phx_source_groups(s, src.id, &ng) != PHX_OK) continue;
        for (int g = 0; g < ng; g++) {
            double re, im;
            int cst;
            phx_source_temporal(s, src.id, g, o.in.t, &re, &im, nullptr, nullptr, nullptr, &cst);
            if (cst) continue;
            const std::string key = std::string(src.name) + "_" + std::to_string(g);
            if (src.complex_value) {
                o.scalars[key + "_real"].push_back(re);
                o.scalars[key + "_imag"].push_back(im);
            } else {
                o.scalars[key].push_back(re);
            }
        }
    }
}

/* header of the k-space matrices (solver_output_matrices.cu:23): LX, LY = -(extent / N), DX, DY = 2 / N */
std::string fft_header(int nc, int nr, double lx, double ly, double dx, double dy, double t) {
    std::ostringstream h;
    h << "# SIZE " << nc << " " << nr << " LX " << lx << " LY " << ly << " DX " << dx << " DY " << dy << " TIME " << t
      << " :: PHOENIX_ MATRIX\n";
    return h.str();
}

/* Solver::outputMatrices (solver_output_matrices.cu:20-70) for wavefunction, reservoir and the k-space spectra */
template <typename real>
void output_matrices(phx_solver* s, const Out& o, int x0, int x1, int y0, int y1, int inc, const std::string& suffix,
                     const std::string& prefix) {
    const char* pm[2] = {"plus", "minus"};
    const int nc = (x1 - x0 + inc - 1) / inc, nr = (y1 - y0 + inc - 1) / inc;
    const double Lx = o.oc.L_x * (x1 - x0) / o.in.N_c, Ly = o.oc.L_y * (y1 - y0) / o.in.N_r;
    const std::string head = header(o, nc, nr, Lx, Ly, o.in.t);
    std::vector<double> z;
    for (int k = 0; k < (o.in.tetm ? 2 : 1); k++) {
        const std::string sfx = pm[k];
        if (o.want({"wavefunction", "psi", k == 0 ? "wavefunction_plus" : "wavefunction_minus", k == 0 ? "psi_plus" : "psi_minus",
                    "plus", "wf", "mat", "all", "initial", k == 0 ? "initial_plus" : "initial_minus"}) &&
            field<real>(s, k == 0 ? PHX_PSI_PLUS : PHX_PSI_MINUS, true, z))
            write_matrix(o, prefix + "wavefunction_" + sfx + suffix, z, x0, x1, y0, y1, inc, head);
        if (o.in.use_reservoir &&
            o.want({"mat", "reservoir", "n", k == 0 ? "reservoir_plus" : "reservoir_minus", k == 0 ? "n_plus" : "n_minus", "plus",
                    "rv", "all"}) &&
            field<real>(s, k == 0 ? PHX_N_PLUS : PHX_N_MINUS, o.in.complex_reservoir != 0, z))
            write_matrix(o, prefix + "reservoir_" + sfx + suffix, z, x0, x1, y0, y1, inc, head);
        /* the spectrum of the last FFT time (matrix.fft_plus / fft_minus) */
        if (o.in.fft_every < o.in.t_max && o.want({"fft_mask", "fft", k == 0 ? "fft_plus" : "fft_minus", "plus", "mat", "all"}) &&
            field<real>(s, k == 0 ? PHX_FFT_PLUS : PHX_FFT_MINUS, true, z))
            write_matrix(o, prefix + "fft_" + sfx + suffix, z, x0, x1, y0, y1, inc,
                         fft_header(nc, nr, -1.0 * (x1 - x0) / o.in.N_c, -1.0 * (y1 - y0) / o.in.N_r, 2.0 / o.in.N_c,
                                    2.0 / o.in.N_r, o.in.t));
    }
}

/* Solver::cacheMatrices (solver_cache.cu:57-72): --historyMatrix snapshots under timeoutput/ */
template <typename real>
void cache_matrices(phx_solver* s, Out& o) {
    if (!o.oc.history || o.in.t < o.oc.hstart) return;
    if (o.history_counter < o.oc.hevery) {
        o.history_counter++;
        return;
    }
    o.history_counter = 1;
    mkdir((o.path + o.prefix + "timeoutput").c_str(), 0755);
    char sfx[64];
    std::snprintf(sfx, sizeof sfx, "_%f", o.in.t); /* std::to_string of the time */
    output_matrices<real>(s, o, o.oc.hx0, o.oc.hx1, o.oc.hy0, o.oc.hy1, o.oc.hinc, sfx, "timeoutput/");
}

/* Solver::cacheToFiles (solver_cache.cu:42-55): scalar.txt, one row per output interval */
void cache_to_files(const Out& o) {
    if (!o.want({"all", "max", "scalar"})) return;
    FILE* f = std::fopen((o.path + o.prefix + "scalar.txt").c_str(), "w");
    if (!f) return;
    std::fprintf(f, "index ");
    for (const auto& kv : o.scalars) std::fprintf(f, "%s ", kv.first.c_str());
    std::fprintf(f, "\n");
    const size_t rows = o.scalars.count("t") ? o.scalars.at("t").size() : 0;
    for (size_t i = 0; i < rows; i++) {
        std::fprintf(f, "%zu ", i);
        for (const auto& kv : o.scalars) std::fprintf(f, "%.10g ", i < kv.second.size() ? kv.second[i] : NAN);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

/* SystemParameters::printCMD (system_print_output.cpp:812-848), without cursor control */
void print_progress(const phx_info& in, double runtime, int64_t iterations) {
    const int width = 50;
    const int filled = in.t_max > 0 ? (int)std::min<double>(width, width * in.t / in.t_max) : width;
    std::string bar = "[" + std::string(filled, '=') + std::string(width - filled, ' ') + "]";
    std::printf("    T = %dps - dt = %.2gps\n", (int)in.t, in.dt);
    std::printf("    Progress: %s %d%%\n", bar.c_str(), in.t_max > 0 ? (int)(100.0 * in.t / in.t_max) : 100);
    std::printf("    Current System: %s - %s\n", in.tetm ? "TE/TM" : "Scalar", in.use_stochastic ? "With Stochastic" : "No Stochastic");
    std::printf("    Runtime: %ds, remaining: %ds\n", (int)runtime, in.t > 0 ? (int)(runtime * (in.t_max - in.t) / in.t) : 0);
    std::printf("    Time per ps: %gs/ps-  %.3gps/s-  %git/s\n", in.t > 0 ? runtime / in.t : 0.0, runtime > 0 ? in.t / runtime : 0.0,
                runtime > 0 ? (double)iterations / runtime : 0.0);
    std::fflush(stdout);
}

template <typename real>
int run(phx_solver* s) {
    Out o;
    phx_get_output_config(s, &o.oc);
    phx_get_info(s, &o.in);
    std::stringstream ks(o.oc.keys);
    for (std::string k; std::getline(ks, k, ',');) o.keys.push_back(k);
    o.ftype = o.oc.filetype[0] ? o.oc.filetype : "txt";
    o.fp32 = sizeof(real) == 4;
    o.path = o.oc.path;
    if (o.path.empty()) o.path = "data/";
    if (o.path.back() != '/') o.path += "/";
    o.prefix = o.oc.name[0] ? std::string(o.oc.name) + "_" : std::string();
    mkdir(o.path.c_str(), 0755);
    output_initial<real>(s, o);

    auto t0 = std::chrono::steady_clock::now();
    auto last_print = t0 - std::chrono::seconds(1);
    int64_t done = 0;
    int finished = 0;
    while (!finished) {
        int64_t d = 0;
        if (phx_run_interval(s, &d, &finished) != PHX_OK) {
            std::fprintf(stderr, "Run failed: %s\n", phx_last_error());
            return 1;
        }
        done += d;
        cache_values<real>(s, o);
        cache_matrices<real>(s, o);
        const auto now = std::chrono::steady_clock::now();
        if (now - last_print >= std::chrono::milliseconds(250) || finished) {
            print_progress(o.in, std::chrono::duration<double>(now - t0).count(), done);
            last_print = now;
        }
    }
    phx_sync(s);
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    phx_get_info(s, &o.in);
    output_matrices<real>(s, o, 0, o.in.N_c, 0, o.in.N_r, 1, "", "");
    cache_to_files(o);
    std::printf("Total Runtime: %g s --> %g ms/ps --> %g ps/s --> %g mus/it\n", total, total / o.in.t * 1E3, o.in.t / total,
                done > 0 ? total / (double)done * 1e6 : 0.0);
    return 0;
}

/* --help / -h, --helpVerbose / -hv, with --markdown / -mk as a table (system.cpp:162-170: print, exit 0) */
struct FlagHelp {
    const char* flags;
    const char* args;
    const char* what;
    bool verbose_only;
};
const FlagHelp kFlagHelp[] = {
    {"--N, --gridsize", "N_c N_r", "grid points (even)", false},
    {"--L, --gridlength, --xmax", "L_x L_y", "domain lengths in um; dx = L / (N - 1)", false},
    {"--boundary", "bx by", "periodic or zero, per axis", false},
    {"--tmax, --tend", "t", "end time in ps", false},
    {"--tstep, --dt", "dt", "time step in ps (default: 0.5 dx dy / (m_e / e 1e8 / m_eff))", false},
    {"--t0", "t", "start time", true},
    {"--iterator", "key", "RK4 (default), DP45, Heun2, BoSha, CashKarp, Fehlberg2/5, NSRK78, DP85, ..., SSFM", false},
    {"-ssfm", "", "the split-step Fourier iterator", true},
    {"-adaptive, --adaptiveTimestep", "", "adaptive step size for iterators that have one", false},
    {"--tol", "tolerance", "adaptive error tolerance", false},
    {"--rkvdt", "dt_min dt_max", "adaptive step bounds", false},
    {"--imagTime", "amplitude", "imaginary-time propagation, fields normalised to sum |f|^2 dV = amplitude", false},
    {"-tetm", "", "two-component (TE-TM) spinor model", false},
    {"-noReservoir", "", "no reservoir even with a pump", true},
    {"--pump, --potential, --pulse", "envelope", "sources (several allowed; see the envelope syntax)", false},
    {"--fftMask", "envelope", "k-space filter mask, applied every --fftEvery", false},
    {"--initialState, --initState, --iS", "envelope", "initial wavefunction", false},
    {"--initialReservoir, --initReservoir, --iR", "envelope", "initial reservoir", true},
    {"--fftEvery", "dt", "interval of the k-space filter (and of the spectra output)", false},
    {"--dw", "amplitude", "stochastic noise amplitude", false},
    {"--initRandom", "amplitude seed", "random initial wavefunction; seed of the noise", false},
    {"--gammaC, --gammaR, --gc, --gr, --R, --gpm, --deltaLT", "value", "model parameters", false},
    {"--meff, --hbarscaled, --hbar, --e, --me", "value", "constants", true},
    {"--path, --outputPath", "dir", "output directory (default data/)", false},
    {"--name", "prefix", "prefix of every output file: <path><prefix>_<file>", false},
    {"--output", "keys", "comma-separated output keys (default mat,scalar)", false},
    {"--outEvery", "dt", "output interval", false},
    {"--outputFileType", "type", "txt (default), npy, ash or mat", false},
    {"--fileoutPrecision", "digits", "significant digits of text output", true},
    {"--historyMatrix", "x0 x1 y0 y1 inc", "region of the matrix snapshots under timeoutput/", true},
    {"--historyTime", "start every", "when snapshots are taken", true},
    {"-verboseIO", "", "report each written matrix", true},
    {"--config", "file", "read further flags from a file", false},
    {"--subgrids, --sg, --threads, --blocksize", "...", "accepted for compatibility; one grid on one GPU here", true},
    {"--help, -h / --helpVerbose, -hv", "", "this list (verbose: every flag and the envelope syntax); --markdown, -mk: as a table",
     false},
};

bool has_flag(int argc, char** argv, std::initializer_list<const char*> names) {
    for (int i = 1; i < argc; i++)
        for (const char* n : names)
            if (std::strcmp(argv[i], n) == 0) return true;
    return false;
}

void print_help(bool verbose, bool markdown) {
    std::printf(markdown ? "# PHOENIX (MI355X) command line\n\n| Flag | Arguments | Meaning |\n|---|---|---|\n"
                         : "PHOENIX (MI355X) command line:\n\n");
    for (const FlagHelp& f : kFlagHelp) {
        if (f.verbose_only && !verbose) continue;
        if (markdown)
            std::printf("| `%s` | %s | %s |\n", f.flags, f.args, f.what);
        else
            std::printf("  %-44s %-18s %s\n", f.flags, f.args, f.what);
    }
    if (verbose)
        std::printf("\nEnvelope syntax: amplitude behaviour width_x width_y pos_x pos_y polarization exponent charge type\n"
                    "[time kind t0 sigma freq], or with named fields after amplitude and behaviour: width w_x w_y, pos x y,\n"
                    "pol p, exponent e, charge m, type t, momenta k_x k_y, time kind t0 sigma freq; \"load file\" reads a\n"
                    "matrix.  behaviour: add, multiply, replace, adaptive, complex; type: gauss, ring, noDivide, local,\n"
                    "outerExponent, joined with '+'; time kinds: constant, cos, gauss, iexp, load.\n");
}

} // namespace

int main(int argc, char** argv) {
    if (has_flag(argc, argv, {"--help", "-h", "--helpVerbose", "-hv"})) {
        print_help(has_flag(argc, argv, {"--helpVerbose", "-hv"}), has_flag(argc, argv, {"--markdown", "-mk"}));
        return 0;
    }
    phx_solver* s = nullptr;
    const int prec = PHX_MAIN_PRECISION;
    if (phx_create(&s, prec, argc, (const char* const*)argv) != PHX_OK) {
        std::fprintf(stderr, "Invalid input! %s\n", phx_last_error());
        return 1;
    }
    if (phx_initialize(s) != PHX_OK) {
        std::fprintf(stderr, "Initialisation failed: %s\n", phx_last_error());
        phx_destroy(s);
        return 1;
    }
    const int rc = prec == PHX_FP32 ? run<float>(s) : run<double>(s);
    phx_destroy(s);
    return rc;
}
