/*
 * phx_config.hpp -- PHOENIX command line / config surface and input envelopes
 * (host C++, no device code).
 *
 * Restates the reference's parsing rules:
 *   - flags with aliases and prefixes ("--" for values, "-" for switches), the
 *     first occurrence wins (commandline_io.cpp:110-126);
 *   - numbers via std::stod, rounded to the build's real type (commandline_io.cpp:129-144);
 *   - --config files append whitespace tokens, '#' starts a comment
 *     (config_readin.cpp:7-42);
 *   - the envelope mini language, positional and named forms (envelope.cpp:148-315);
 *   - defaults and derived constants (system.cpp:18-160);
 *   - validation (system_validate_input.cpp:7-56), reported as an error instead of exit(1).
 */
#pragma once
#include <complex>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "custom_rhs.h"

namespace phx {

/* Envelope enums carry the reference's bit values (envelope.hpp:37-68). */
enum EnvType : uint32_t { ENV_GAUSS = 1, ENV_OUTER_EXPONENT = 2, ENV_RING = 4, ENV_NO_DIVIDE = 8, ENV_LOCAL = 16 };
enum EnvPol : uint32_t { POL_PLUS = 1, POL_MINUS = 2, POL_BOTH = 3 };
enum EnvBehavior : uint32_t { BEH_ADD = 1, BEH_MULTIPLY = 2, BEH_REPLACE = 4, BEH_ADAPTIVE = 8, BEH_COMPLEX = 16 };
enum EnvTemporal : uint32_t { TMP_IEXP = 1, TMP_COS = 2, TMP_GAUSS = 4, TMP_CONSTANT = 8, TMP_LOADED = 16 };

struct Envelope {
    struct Spatial {
        double amp = 0, width_x = 0, width_y = 0, x = 0, y = 0, exponent = 0, k0_x = 0, k0_y = 0;
        uint32_t type = ENV_GAUSS, pol = POL_BOTH, behavior = BEH_ADD;
        int m = 0;
        std::string load_path;
        int group = 0;
    };
    struct Temporal {
        uint32_t kind = TMP_CONSTANT;
        double t0 = 0, sigma = 0, freq = 0;
        std::string load_path;
        std::vector<double> tp, vr, vi; /* loaded points */
    };
    std::vector<Spatial> comps;
    std::vector<Temporal> groups;
    std::map<std::string, int> group_ids;

    size_t size() const { return comps.size(); }
    int groupSize() const { return (int)groups.size(); }
    bool allConstant() const;
    /* envelope.cpp:466-499 */
    std::complex<double> temporal(int g, double t) const;
    /* envelope.cpp:332-412; group < 0 = all groups.  Returns "" or an error. */
    std::string calculate(std::complex<double>* out, int group, uint32_t pol, int N_c, int N_r, double L_x, double L_y,
                          double dx, double dy, double default_if_unset) const;
};

struct SystemConfig {
    /* SI units and model (system.cpp:20-37) */
    double m_e = 9.10938356E-31, h_bar = 1.0545718E-34, e_e = 1.60217662E-19;
    double h_bar_s = -1, m_eff = -1;
    double gamma_c = 0.15, gamma_r = 1.5 * 0.15, g_c = 3.E-6, g_r = 2. * 3.E-6, R = 0.01;
    double L_x = 100., L_y = 100., g_pm = -3.E-6 / 5, delta_LT = 0.025E-3;
    /* numerics (system.cpp:39-62) */
    int N_c = 400, N_r = 400, subgrids_columns = 0, subgrids_rows = 0;
    double t_max = 1000, t = 0, dt = 0, output_every = 1, dt_max = 3, dt_min = 0.0001, tolerance = 1E-1;
    bool do_overwrite_dt = true;
    double fft_every = 1;
    int block_size = 256, threads = 4;
    std::string iterator = "RK4";
    bool adaptive = false, tetm = false;
    bool random_init = false;
    double random_amp = 1.0;
    uint32_t random_seed = 0;
    double stochastic_amplitude = 0.0;
    double imag_time_amplitude = 0.0;
    bool periodic_x = false, periodic_y = false;
    bool no_reservoir_flag = false;
    Envelope pump, potential, pulse, fft_mask, initial_state, initial_reservoir;
    bool use_reservoir = false, use_pumps = false, use_pulses = false, use_potentials = false, use_fft_mask = false,
         use_stochastic = false;
    /* complex reservoir: n held as {re, im} per cell.  The reference stores n complex throughout; it
       becomes complex in imaginary time (n advances by -i dt like psi) and under a complex temporal
       pump amplitude.  Set by finalize_config for those cases; served by the staged path. */
    bool complex_reservoir = false;
    std::string path = "data/";
    std::vector<std::string> output_keys = {"mat", "scalar"};
    /* --historyMatrix x0 x1 y0 y1 increment, --historyTime start every, --fileoutPrecision
       (system_initialization.cpp:160-181) */
    bool history = false;
    int history_x0 = 0, history_x1 = -1, history_y0 = 0, history_y1 = -1, history_increment = 1, history_every = 1;
    double history_start = 0.0;
    int fileout_precision = -1;
    std::string output_filetype = "txt"; /* --outputFileType txt|npy|ash|mat (filehandler.cpp:49-69) */
    std::string output_name;             /* --name: prefix "<name>_" of every output file (filehandler.cpp:40-41,461-463) */
    bool verbose_io = false;             /* -verboseIO: one line per written matrix (filehandler.cpp:72-74) */
    /* custom parameters (custom_rhs.h) */
#define PHX_X(name, def) double name = def;
    PHX_CUSTOM_PARAMETERS(PHX_X)
#undef PHX_X
    /* derived (system.cpp:116-160) */
    double dx = 0, dy = 0, dV = 0, one_over_dx2 = 0, one_over_dy2 = 0, m2_over_dx2_p_dy2 = 0, m_eff_scaled = 0,
           one_over_h_bar_s = 0, magic_timestep = 0;
    int halo = 4;
    bool fp32 = false;
};

/* Parse argv tokens (argv[0] = program name).  Returns "" on success. */
std::string parse_config(SystemConfig& c, const std::vector<std::string>& argv_in, bool fp32);

/* calculateAuto + validateInputs; returns "" or the validation failure. */
std::string finalize_config(SystemConfig& c);

/* Device-facing physics constants in the build's precision. */
template <typename real>
PhxPhys<real> make_phys(const SystemConfig& c);

/* PHOENIX txt matrix loader (filehandler.cpp:160-198): header "# SIZE N_c N_r ...",
   N real parts then N imaginary parts.  Returns false if the file cannot be opened. */
bool load_txt_matrix(const std::string& path, std::vector<std::complex<double>>& out, int& N_c, int& N_r);

} // namespace phx
