/*
 * phx_main.cpp -- PHOENIX_gpu_{fp32,fp64}: command line front end over libphx.
 *
 * Mirrors main() (source/main.cu:46-135): parse flags, initialise, run until
 * t_max with the output landing of the main loop, write the final matrices
 * (solver_output_matrices.cu:20-70, txt format of filehandler.cpp:282-307) and
 * print the runtime summary line that tooling parses
 * (system_print_output.cpp:731, "... mus/it").  Errors exit with status 1
 * (system_validate_input.cpp:53-56).
 */
#include <chrono>
#include <cstdio>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "phx.h"

#ifndef PHX_MAIN_PRECISION
#define PHX_MAIN_PRECISION 1
#endif

template <typename real>
static void write_txt(const std::string& file, const std::vector<real>& buf, bool is_complex, int N_c, int N_r, const phx_info& in, double L_x, double L_y) {
    FILE* f = std::fopen(file.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "Unable to write '%s'\n", file.c_str());
        return;
    }
    std::fprintf(f, "# SIZE %d %d LX %g LY %g DX %g DY %g TIME %g :: PHOENIX_ MATRIX\n", N_c, N_r, L_x, L_y, in.dx, in.dy, in.t);
    for (int part = 0; part < 2; part++) {
        for (int r = 0; r < N_r; r++) {
            for (int c = 0; c < N_c; c++) {
                const size_t i = (size_t)r * N_c + c;
                const double v = is_complex ? (double)buf[2 * i + part] : (part == 0 ? (double)buf[i] : 0.0);
                std::fprintf(f, "%.17g ", v);
            }
            std::fprintf(f, "\n");
        }
    }
    std::fclose(f);
}

static std::string path_of(int argc, char** argv) {
    std::string p = "data/";
    for (int i = 1; i + 1 < argc; i++)
        if (std::string(argv[i]) == "--path") {
            p = argv[i + 1];
            break;
        }
    if (!p.empty() && p.back() != '/') p += "/";
    return p;
}

template <typename real>
static int finalize(phx_solver* s, const std::string& path) {
    phx_info in;
    phx_get_info(s, &in);
    mkdir(path.c_str(), 0755);
    const size_t N = (size_t)in.N_c * in.N_r;
    const double L_x = in.dx * (in.N_c - 1), L_y = in.dy * (in.N_r - 1);
    std::vector<real> psi(2 * N), n(N);
    const char* names[2] = {"plus", "minus"};
    for (int k = 0; k < (in.tetm ? 2 : 1); k++) {
        if (phx_get_field(s, k == 0 ? PHX_PSI_PLUS : PHX_PSI_MINUS, psi.data(), psi.size() * sizeof(real)) == PHX_OK)
            write_txt(path + "wavefunction_" + names[k] + ".txt", psi, true, in.N_c, in.N_r, in, L_x, L_y);
        if (in.use_reservoir && phx_get_field(s, k == 0 ? PHX_N_PLUS : PHX_N_MINUS, n.data(), n.size() * sizeof(real)) == PHX_OK)
            write_txt(path + "reservoir_" + names[k] + ".txt", n, false, in.N_c, in.N_r, in, L_x, L_y);
    }
    return 0;
}

int main(int argc, char** argv) {
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
    auto t0 = std::chrono::steady_clock::now();
    int64_t done = 0;
    int rc = phx_run(s, &done);
    phx_sync(s);
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != PHX_OK) {
        std::fprintf(stderr, "Run failed: %s\n", phx_last_error());
        phx_destroy(s);
        return 1;
    }
    phx_info in;
    phx_get_info(s, &in);
    const std::string path = path_of(argc, argv);
    if (prec == PHX_FP32)
        finalize<float>(s, path);
    else
        finalize<double>(s, path);
    std::printf("Total Runtime: %g s --> %g ms/ps --> %g ps/s --> %g mus/it\n", total, total / in.t * 1E3, in.t / total,
                done > 0 ? total / (double)done * 1e6 : 0.0);
    phx_destroy(s);
    return 0;
}
