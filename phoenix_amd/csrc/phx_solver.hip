/*
 * phx_solver.hip -- the solver object behind the C ABI of include/phx.h.
 *
 * Owns the device fields (full grid, no halos: the fused kernels resolve the
 * boundaries), drives iterate() with the reference's time bookkeeping
 * (solver.cu:22-70) and the main-loop output landing (main.cu:80-117), and
 * evaluates the envelopes on the host once (solver_initialization.cu:37-174).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/phx.h"
#include "phx_config.hpp"
#include "phx_kernels.h"

namespace {
thread_local std::string g_err;
}

namespace phx {
/* shared by every entry point of the library: phx_last_error() reports it */
int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
} // namespace phx

namespace {
int fail(int code, const std::string& msg) { return phx::set_error(code, msg); }

#define PHX_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) return fail(PHX_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

/* PHX_RK4_TY / PHX_RK4_EW / PHX_RK4_PF override the launch shape (used by the tuning sweep). */
phx::Rk4Tune rk4_tune_from_env(phx::Rk4Tune t) {
    if (const char* e = std::getenv("PHX_RK4_TY")) t.ty = std::atoi(e);
    if (const char* e = std::getenv("PHX_RK4_EW")) t.ew = std::atoi(e);
    if (const char* e = std::getenv("PHX_RK4_PF")) t.pf = std::atoi(e);
    return t;
}
} // namespace

struct phx_solver {
    phx::SystemConfig cfg;
    int precision = PHX_FP64;
    size_t rsize = 8; /* sizeof(real) */
    hipStream_t own_stream = nullptr, stream = nullptr;
    bool initialized = false;
    /* the kernels' physics constants, derived once per solver (they do not change after phx_create) */
    PhxPhys<float> ph_f{};
    PhxPhys<double> ph_d{};
    /* device fields: [component][buffer] */
    void* psi[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    void* n[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int cur = 0; /* current buffer */
    void* pump[2] = {nullptr, nullptr};
    void* pot[2] = {nullptr, nullptr};
    void* mask[2] = {nullptr, nullptr};
    void* tw_c = nullptr; /* forward FFT twiddles of the row length N_c */
    void* tw_r = nullptr; /* ... and of the column length N_r */
    void* noise = nullptr;      /* noise field of the current step (complex, real precision) */
    double* norm_partials = nullptr; /* imaginary-time normalisation: per-block sums, reduced sums */
    double* norm_sums = nullptr;
    double* partials = nullptr; /* per-wavefront (error, norm) of the adaptive kernel */
    double* sums = nullptr;     /* reduced (error, norm), device */
    double* sums_host = nullptr; /* pinned host copy */
    const phx::IteratorInfo* iter = nullptr; /* registry entry of cfg.iterator (nullptr for SSFM) */
    bool ssfm = false;                        /* split-step Fourier iterator */
    void* ssfm_rowf = nullptr;                /* linear half-step factors along rows (N_r) and columns (N_c) */
    void* ssfm_colf = nullptr;
    double ssfm_dt = -1.0;                    /* dt the factors were made for */
    int max_rejections = 10000;
    double last_error = 0; /* E of the last adaptive attempt */
    /* time bookkeeping in the build's precision (system_parameters.hpp:33) */
    double t = 0, dt = 0, fft_cached_t = 0;
    int64_t iteration = 0, rejected = 0;
    int64_t out_iter = 1; /* output intervals started (main.cu:64, out_every_iterations) */
    phx::Rk4Tune tune; /* launch shape of the fused RK4 kernel */
    bool tuned = false;
    /* captured run of kGraphSteps fused RK4 launches (replayed while nothing changes between steps) */
    hipGraphExec_t graph = nullptr;
    int graph_parity = -1;
    double graph_dt = 0;
    hipStream_t graph_stream = nullptr;
    bool graph_off = false;
    /* the same runs as one launch of rk4_scalar_flow: per-tile step counters (all equal to
       flow_base between launches) and the kernel's error word */
    int* flow_counters = nullptr;
    int* flow_error = nullptr;
    int flow_base = 0;
    int64_t flow_steps = 0, flow_launches = 0;
    bool flow_off = false;
    /* staged path (phx_staged.hip): padded stage input and K slots, per-group source fields */
    bool staged = false;
    bool grouped = false; /* one source field per temporal group (staged path, SSFM with chains or pulses) */
    /* k-space spectra kept for output (the reference's fft_plus / fft_minus): the filtered spectrum
       of the last filter application, or, without a mask, a copy of psi at the last FFT time whose
       forward transform is taken when the spectrum is read (spec_raw) */
    void* spec[2] = {nullptr, nullptr};
    bool spec_raw = false;
    int staged_H = 0;
    size_t staged_padN = 0;
    void* st_y_psi[2] = {nullptr, nullptr};
    void* st_y_n[2] = {nullptr, nullptr};
    void* st_k_psi[2] = {nullptr, nullptr};
    void* st_k_n[2] = {nullptr, nullptr};
    void* st_pulse[2] = {nullptr, nullptr};
    int st_npump = 0, st_npot = 0, st_npulse = 0;
    double* st_partials = nullptr;
    std::vector<void*> allocs;

    size_t cells() const { return (size_t)cfg.N_c * cfg.N_r; }
    double rnd(double x) const { return precision == PHX_FP32 ? (double)(float)x : x; }
};

namespace {

int dev_alloc(phx_solver* s, void** p, size_t bytes) {
    PHX_HIP(hipMalloc(p, bytes));
    PHX_HIP(hipMemsetAsync(*p, 0, bytes, s->stream));
    s->allocs.push_back(*p);
    return PHX_OK;
}

template <typename real>
int upload_real(phx_solver* s, void* dst, const std::vector<std::complex<double>>& src) {
    std::vector<real> h(src.size());
    for (size_t i = 0; i < src.size(); i++) h[i] = (real)src[i].real();
    PHX_HIP(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(real), hipMemcpyHostToDevice, s->stream));
    PHX_HIP(hipStreamSynchronize(s->stream));
    return PHX_OK;
}

template <typename real>
int upload_complex(phx_solver* s, void* dst, const std::vector<std::complex<double>>& src) {
    std::vector<real> h(2 * src.size());
    for (size_t i = 0; i < src.size(); i++) {
        h[2 * i] = (real)src[i].real();
        h[2 * i + 1] = (real)src[i].imag();
    }
    PHX_HIP(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(real), hipMemcpyHostToDevice, s->stream));
    PHX_HIP(hipStreamSynchronize(s->stream));
    return PHX_OK;
}

/* Sum of all spatial groups, each at temporal amplitude 1 (valid when all groups are
   constant) or a single group whose amplitude is applied per stage. */
int env_field(phx_solver* s, const phx::Envelope& e, uint32_t pol, std::vector<std::complex<double>>& out) {
    const auto& c = s->cfg;
    out.assign(s->cells(), {0.0, 0.0});
    std::vector<std::complex<double>> tmp(s->cells());
    for (int g = 0; g < e.groupSize(); g++) {
        e.calculate(tmp.data(), g, pol, c.N_c, c.N_r, c.L_x, c.L_y, c.dx, c.dy, 0.0);
        for (size_t i = 0; i < tmp.size(); i++) out[i] += tmp[i];
    }
    return PHX_OK;
}

template <typename real>
const PhxPhys<real>& phys_of(const phx_solver* s) {
    if constexpr (sizeof(real) == 4)
        return s->ph_f;
    else
        return s->ph_d;
}

template <typename real>
int initialize_t(phx_solver* s) {
    auto& c = s->cfg;
    s->ph_f = phx::make_phys<float>(c);
    s->ph_d = phx::make_phys<double>(c);
    const size_t N = s->cells();
    const int ncomp = c.tetm ? 2 : 1;
    int n_it = 0;
    const phx::IteratorInfo* table = phx::iterator_table(&n_it);
    for (int k = 0; k < n_it && !s->iter; k++)
        if (c.iterator == table[k].key) s->iter = &table[k];
    s->ssfm = c.iterator == "SSFM";
    if (!s->iter && !s->ssfm) return fail(PHX_ERR_UNSUPPORTED, "iterator '" + c.iterator + "' is not provided by this library");
    if (s->ssfm && (!phx::fft_supported(c.N_c, sizeof(real)) || !phx::fft_supported(c.N_r, sizeof(real))))
        return fail(PHX_ERR_UNSUPPORTED, std::string("SSFM needs grid sizes of at most ") +
                                             (sizeof(real) == 4 ? "8192 (powers of two) or 4096" : "4096 (powers of two) or 2048") +
                                             " points per axis");
    if (c.adaptive && s->iter && s->iter->can_adapt) {
        const int nw = phx::rk_generic_waves(s->iter->adaptive, c.N_c, c.N_r);
        PHX_HIP(hipMalloc((void**)&s->partials, 2 * sizeof(double) * (size_t)nw));
        s->allocs.push_back(s->partials);
        PHX_HIP(hipMalloc((void**)&s->sums, 2 * sizeof(double)));
        s->allocs.push_back(s->sums);
        PHX_HIP(hipHostMalloc((void**)&s->sums_host, 2 * sizeof(double), hipHostMallocDefault));
    }
    if (const char* e = std::getenv("PHX_MAX_REJECTIONS")) s->max_rejections = std::atoi(e);
    if (c.use_fft_mask && (!phx::fft_supported(c.N_c, sizeof(real)) || !phx::fft_supported(c.N_r, sizeof(real))))
        return fail(PHX_ERR_UNSUPPORTED, std::string("--fftMask needs grid sizes of at most ") +
                                             (sizeof(real) == 4 ? "8192 (powers of two) or 4096" : "4096 (powers of two) or 2048") +
                                             " points per axis");
    /* The staged path serves what the fused kernels do not: more than seven stages, pulses,
       several non-constant temporal groups of a source, complex temporal potential envelopes, the
       complex reservoir. */
    const size_t nel = c.complex_reservoir ? 2 : 1; /* reals per reservoir cell */
    const auto complex_temporal = [](const phx::Envelope& e) {
        for (int g = 0; g < e.groupSize(); g++)
            if (e.groups[g].kind & phx::TMP_IEXP) return true;
        return false;
    };
    s->grouped = c.use_pulses || (c.use_pumps && c.pump.groupSize() > 1 && !c.pump.allConstant()) ||
                 (c.use_potentials && c.potential.groupSize() > 1 && !c.potential.allConstant()) ||
                 (c.use_potentials && complex_temporal(c.potential));
    /* SSFM takes the group fields in its own pointwise kernels (phx_ssfm.hip) */
    s->staged = !s->ssfm && (s->grouped || c.complex_reservoir ||
                             (s->iter && (!phx::rk_tableau_fused(s->iter->fixed) || !phx::rk_tableau_fused(s->iter->adaptive))));
    if (s->staged) s->grouped = true;

    for (int k = 0; k < ncomp; k++) {
        for (int b = 0; b < 2; b++) {
            if (int r = dev_alloc(s, &s->psi[k][b], 2 * N * sizeof(real))) return r;
            if (c.use_reservoir)
                if (int r = dev_alloc(s, &s->n[k][b], nel * N * sizeof(real))) return r;
        }
        /* staged path: one field per temporal group (group g at g N); the fused kernels: one
           field, all groups summed (constant envelopes) or the single group */
        s->st_npump = (s->grouped && c.use_reservoir && c.use_pumps) ? c.pump.groupSize() : 0;
        s->st_npot = (s->grouped && c.use_potentials) ? c.potential.groupSize() : 0;
        s->st_npulse = (s->grouped && c.use_pulses) ? c.pulse.groupSize() : 0;
        if (s->st_npump > phx::kStagedMaxGroups || s->st_npot > phx::kStagedMaxGroups || s->st_npulse > phx::kStagedMaxGroups)
            return fail(PHX_ERR_UNSUPPORTED, "more than " + std::to_string(phx::kStagedMaxGroups) + " temporal groups of one source");
        if (c.use_reservoir && c.use_pumps)
            if (int r = dev_alloc(s, &s->pump[k], std::max(1, s->st_npump) * N * sizeof(real))) return r;
        if (c.use_potentials)
            if (int r = dev_alloc(s, &s->pot[k], std::max(1, s->st_npot) * N * sizeof(real))) return r;
        if (s->st_npulse)
            if (int r = dev_alloc(s, &s->st_pulse[k], s->st_npulse * 2 * N * sizeof(real))) return r;
        if (c.use_fft_mask)
            if (int r = dev_alloc(s, &s->mask[k], N * sizeof(real))) return r;
    }
    if (c.use_stochastic)
        if (int r = dev_alloc(s, &s->noise, 2 * N * sizeof(real))) return r;
    /* fft_plus / fft_minus are written when fft_every < t_max and the output keys ask for them
       (solver_output_matrices.cu:41,62) */
    const bool fft_sizes = phx::fft_supported(c.N_c, sizeof(real)) && phx::fft_supported(c.N_r, sizeof(real));
    for (int k = 0; k < ncomp && c.fft_every < c.t_max && fft_sizes; k++) {
        const char* keys[] = {"fft_mask", "fft", k == 0 ? "fft_plus" : "fft_minus", "plus", "mat", "all"};
        bool want = false;
        for (const char* key : keys) want = want || std::find(c.output_keys.begin(), c.output_keys.end(), key) != c.output_keys.end();
        if (want)
            if (int r = dev_alloc(s, &s->spec[k], 2 * N * sizeof(real))) return r;
    }
    if (s->staged) {
        const int H = std::max(phx::rk_tableau_stages(s->iter->fixed), phx::rk_tableau_stages(s->iter->adaptive));
        s->staged_H = H;
        s->staged_padN = (size_t)(c.N_c + 2 * H) * (size_t)(c.N_r + 2 * H);
        const size_t pn = s->staged_padN;
        for (int k = 0; k < ncomp; k++) {
            if (int r = dev_alloc(s, &s->st_y_psi[k], 2 * pn * sizeof(real))) return r;
            if (int r = dev_alloc(s, &s->st_k_psi[k], (size_t)H * 2 * pn * sizeof(real))) return r;
            if (c.use_reservoir) {
                if (int r = dev_alloc(s, &s->st_y_n[k], nel * pn * sizeof(real))) return r;
                if (int r = dev_alloc(s, &s->st_k_n[k], (size_t)H * nel * pn * sizeof(real))) return r;
            }
        }
        const size_t nb = phx::staged_final_blocks(c.N_c, c.N_r);
        if (int r = dev_alloc(s, (void**)&s->st_partials, 2 * nb * sizeof(double))) return r;
        if (!s->sums) {
            if (int r = dev_alloc(s, (void**)&s->sums, 2 * sizeof(double))) return r;
            PHX_HIP(hipHostMalloc((void**)&s->sums_host, 2 * sizeof(double), hipHostMallocDefault));
        }
    }
    if (c.imag_time_amplitude != 0.0) {
        if (int r = dev_alloc(s, (void**)&s->norm_partials, 2 * sizeof(double) * phx::imag_norm_blocks(N))) return r;
        if (int r = dev_alloc(s, (void**)&s->norm_sums, 2 * sizeof(double))) return r;
    }
    if (s->ssfm) {
        if (int r = dev_alloc(s, &s->ssfm_rowf, 2 * (size_t)c.N_r * sizeof(real))) return r;
        if (int r = dev_alloc(s, &s->ssfm_colf, 2 * (size_t)c.N_c * sizeof(real))) return r;
    }
    if (c.use_fft_mask || s->ssfm || s->spec[0]) {
        for (int d = 0; d < 2; d++) {
            const int L = d == 0 ? c.N_c : c.N_r;
            std::vector<std::complex<double>> w(L);
            for (int k = 0; k < L; k++) w[k] = std::polar(1.0, -2.0 * M_PI * (double)k / (double)L);
            void** slot = d == 0 ? &s->tw_c : &s->tw_r;
            if (int r = dev_alloc(s, slot, 2 * L * sizeof(real))) return r;
            if (int r = upload_complex<real>(s, *slot, w)) return r;
        }
    }

    /* initial states (solver_initialization.cu:55-98) */
    std::vector<std::complex<double>> psi0[2], n0[2];
    for (int k = 0; k < ncomp; k++) {
        const uint32_t pol = k == 0 ? phx::POL_PLUS : phx::POL_MINUS;
        psi0[k].assign(N, {0.0, 0.0});
        c.initial_state.calculate(psi0[k].data(), -1, pol, c.N_c, c.N_r, c.L_x, c.L_y, c.dx, c.dy, 0.0);
        n0[k].assign(N, {0.0, 0.0});
        if (c.use_reservoir) c.initial_reservoir.calculate(n0[k].data(), -1, pol, c.N_c, c.N_r, c.L_x, c.L_y, c.dx, c.dy, 0.0);
    }
    if (c.random_init) {
        /* std::mt19937 + uniform_real_distribution<real>(-a, a), psi+ then psi-, row-major */
        std::mt19937 gen{c.random_seed};
        std::uniform_real_distribution<real> dist{(real)-c.random_amp, (real)c.random_amp};
        for (int k = 0; k < ncomp; k++)
            for (auto& z : psi0[k]) {
                real a = dist(gen);
                real b = dist(gen);
                z += std::complex<double>((double)a, (double)b);
            }
    }
    for (int k = 0; k < ncomp; k++) {
        if (int r = upload_complex<real>(s, s->psi[k][0], psi0[k])) return r;
        if (c.use_reservoir) {
            /* the reference's reservoir is complex-stored; a real one keeps the real part */
            if (int r = c.complex_reservoir ? upload_complex<real>(s, s->n[k][0], n0[k]) : upload_real<real>(s, s->n[k][0], n0[k]))
                return r;
        }
        const uint32_t pol = k == 0 ? phx::POL_PLUS : phx::POL_MINUS;
        std::vector<std::complex<double>> f;
        /* one spatial field per temporal group (solver_initialization.cu:120-150) */
        auto group_field = [&](const phx::Envelope& e, int g) {
            f.assign(N, {0.0, 0.0});
            e.calculate(f.data(), g, pol, c.N_c, c.N_r, c.L_x, c.L_y, c.dx, c.dy, 0.0);
        };
        if (s->pump[k]) {
            if (s->st_npump) {
                for (int g = 0; g < s->st_npump; g++) {
                    group_field(c.pump, g);
                    if (int r = upload_real<real>(s, (real*)s->pump[k] + (size_t)g * N, f)) return r;
                }
            } else {
                env_field(s, c.pump, pol, f);
                if (int r = upload_real<real>(s, s->pump[k], f)) return r;
            }
        }
        if (s->pot[k]) {
            if (s->st_npot) {
                for (int g = 0; g < s->st_npot; g++) {
                    group_field(c.potential, g);
                    if (int r = upload_real<real>(s, (real*)s->pot[k] + (size_t)g * N, f)) return r;
                }
            } else {
                env_field(s, c.potential, pol, f);
                if (int r = upload_real<real>(s, s->pot[k], f)) return r;
            }
        }
        for (int g = 0; g < s->st_npulse; g++) {
            group_field(c.pulse, g);
            if (int r = upload_complex<real>(s, (real*)s->st_pulse[k] + (size_t)g * 2 * N, f)) return r;
        }
        if (s->mask[k]) {
            f.assign(N, {0.0, 0.0});
            c.fft_mask.calculate(f.data(), -1, pol, c.N_c, c.N_r, c.L_x, c.L_y, c.dx, c.dy, 1.0);
            if (int r = upload_real<real>(s, s->mask[k], f)) return r;
        }
    }
    s->t = s->rnd(c.t);
    s->dt = s->rnd(c.dt);
    s->fft_cached_t = 0.0;
    s->cur = 0;
    s->initialized = true;
    s->tune = rk4_tune_from_env(s->tune);
    return PHX_OK;
}

/* temporal amplitude of a single-group envelope, interpolated as the kernels do
   (kernel_gp_compute.cuh:79): a(t)(1-w) + w a(t+dt) */
template <typename real>
void stage_amps(const phx::Envelope& e, double t, double dt, const double* w, int ns, real* out) {
    real a0 = 1, a1 = 1;
    if (e.groupSize() == 1 && !e.allConstant()) {
        a0 = (real)e.temporal(0, t).real();
        a1 = (real)e.temporal(0, t + dt).real();
    }
    for (int i = 0; i < ns; i++) out[i] = a0 * (real(1.0) - (real)w[i]) + (real)w[i] * a1;
}

template <typename real>
int rk4_launch(phx_solver* s, double t_pre, double dt) {
    auto& c = s->cfg;
    phx::Rk4Args<real> a{};
    using v2 = typename phx::Vec2<real>::type;
    const int nxt = s->cur ^ 1;
    a.psi_in = (const v2*)s->psi[0][s->cur];
    a.psi_out = (v2*)s->psi[0][nxt];
    a.n_in = c.use_reservoir ? (const real*)s->n[0][s->cur] : nullptr;
    a.n_out = c.use_reservoir ? (real*)s->n[0][nxt] : nullptr;
    a.pump = (const real*)s->pump[0];
    a.pot = (const real*)s->pot[0];
    a.Nc = c.N_c;
    a.Nr = c.N_r;
    a.periodic_x = c.periodic_x;
    a.periodic_y = c.periodic_y;
    a.alt_rows = phx::rk4_alt_rows();
    a.dt = (real)dt;
    static const double w[4] = {0.0, 0.5, 0.5, 1.0};
    stage_amps<real>(c.pump, t_pre, dt, w, 4, a.pamp);
    stage_amps<real>(c.potential, t_pre, dt, w, 4, a.vamp_h);
    for (int i = 0; i < 4; i++) a.vamp_h[i] = (real)((double)a.vamp_h[i] * c.one_over_h_bar_s);
    a.ph = phys_of<real>(s);
    PHX_HIP(phx::launch_rk4_scalar<real>(a, s->tune, s->stream));
    s->cur = nxt;
    return PHX_OK;
}

/* Solver::adaptTimeStep (solver.cu:72-105) on FP64 sums over the grid:
   E = |err / norm|, dh = (tol / 2 / max(E, DBL_MIN))^power, NaN -> dh = 0.9 and reject,
   dt' = max(min(dt dh, dt_max), dt_min), accept iff E < tol.  power: 1/4 for DP45,
   CashKarp, Fehlberg2 and Fehlberg5, 1/3 for BoSha (the iterators' adaptTimeStep calls). */
bool adapt_dt(double err, double norm, double tol, double dt, double dt_min, double dt_max, double power, double* new_dt,
              double* final_error) {
    double fe = std::fabs(err / norm);
    double dh = std::pow(tol / 2.0 / std::max(fe, DBL_MIN), power);
    if (std::isnan(dh) || std::isnan(fe)) {
        dh = 0.9;
        fe = DBL_MAX;
    }
    double nd = std::min(dt * dh, dt_max);
    if (nd < dt_min) nd = dt_min;
    *new_dt = nd;
    *final_error = fe;
    return fe < tol;
}

/* One fused generic-tableau step from the current buffers into the other ones. */
template <typename real>
int generic_launch(phx_solver* s, phx::RkTableau tab, double t_k, double dt) {
    auto& c = s->cfg;
    using v2 = typename phx::Vec2<real>::type;
    phx::RkGenArgs<real> a{};
    const int nxt = s->cur ^ 1;
    for (int k = 0; k < (c.tetm ? 2 : 1); k++) {
        a.psi_in[k] = (const v2*)s->psi[k][s->cur];
        a.psi_out[k] = (v2*)s->psi[k][nxt];
        a.n_in[k] = c.use_reservoir ? (const real*)s->n[k][s->cur] : nullptr;
        a.n_out[k] = c.use_reservoir ? (real*)s->n[k][nxt] : nullptr;
        a.pump[k] = (const real*)s->pump[k];
        a.pot[k] = (const real*)s->pot[k];
    }
    a.Nc = c.N_c;
    a.Nr = c.N_r;
    a.periodic_x = c.periodic_x;
    a.periodic_y = c.periodic_y;
    a.tetm = c.tetm;
    a.stochastic = c.use_stochastic;
    a.imag = c.imag_time_amplitude != 0.0;
    a.dt = (real)dt;
    static const double w01[2] = {0.0, 1.0};
    real pa[2], va[2];
    stage_amps<real>(c.pump, t_k, dt, w01, 2, pa);
    stage_amps<real>(c.potential, t_k, dt, w01, 2, va);
    a.amp_pump = pa[0];
    a.amp_next_pump = pa[1];
    a.amp_pot_h = (real)((double)va[0] * c.one_over_h_bar_s);
    a.amp_next_pot_h = (real)((double)va[1] * c.one_over_h_bar_s);
    a.noise = (const v2*)s->noise;
    a.partials = s->partials;
    a.sums = s->sums;
    a.ph = phys_of<real>(s);
    PHX_HIP(phx::launch_rk_generic<real>(a, tab, s->stream));
    return PHX_OK;
}

/* One staged step attempt (phx_staged.hip) from the current buffers into the other ones:
   the reference's CALCULATE_K / INTERMEDIATE_SUM_K sequence on shrinking extents, then the
   final sum (and, for an adaptive tableau, the reduced error and norm in s->sums).  Each
   temporal group g of a source enters stage s with a_g(t)(1 - c_s) + c_s a_g(t + dt)
   (kernel_gp_compute.cuh:79,93,98). */
template <typename real>
int staged_attempt(phx_solver* s, phx::RkTableau tab_id, double t_k, double dt) {
    auto& c = s->cfg;
    using v2 = typename phx::Vec2<real>::type;
    const phx::StagedTab tab = phx::staged_tableau(tab_id);
    phx::StagedArgs<real> a{};
    const int nxt = s->cur ^ 1;
    const int ncomp = c.tetm ? 2 : 1;
    for (int k = 0; k < ncomp; k++) {
        a.x_psi[k] = (const v2*)s->psi[k][s->cur];
        a.out_psi[k] = (v2*)s->psi[k][nxt];
        a.y_psi[k] = (v2*)s->st_y_psi[k];
        a.k_psi[k] = (v2*)s->st_k_psi[k];
        if (c.use_reservoir) {
            a.x_n[k] = (const real*)s->n[k][s->cur];
            a.out_n[k] = (real*)s->n[k][nxt];
            a.y_n[k] = (real*)s->st_y_n[k];
            a.k_n[k] = (real*)s->st_k_n[k];
        }
        a.pump[k] = (const real*)s->pump[k];
        a.pot[k] = (const real*)s->pot[k];
        a.pulse[k] = (const v2*)s->st_pulse[k];
    }
    a.npump = s->st_npump;
    a.npot = s->st_npot;
    a.npulse = s->st_npulse;
    a.noise = c.use_stochastic ? (const v2*)s->noise : nullptr;
    a.stochastic = c.use_stochastic;
    a.tetm = c.tetm;
    a.cres = c.complex_reservoir;
    a.Nc = c.N_c;
    a.Nr = c.N_r;
    a.H = s->staged_H;
    a.P = c.N_c + 2 * s->staged_H;
    a.N = s->cells();
    a.padN = s->staged_padN;
    a.periodic_x = c.periodic_x;
    a.periodic_y = c.periodic_y;
    a.dt = (real)dt;
    const bool imag = c.imag_time_amplitude != 0.0;
    a.dtr = imag ? real(0) : (real)dt;
    a.dti = imag ? -(real)dt : real(0);
    a.partials = s->st_partials;
    a.ph = phys_of<real>(s);
    /* temporal amplitudes of every group at t and t + dt, in the build's precision */
    std::complex<real> amp0[3][phx::kStagedMaxGroups], amp1[3][phx::kStagedMaxGroups];
    const phx::Envelope* env[3] = {&c.pump, &c.potential, &c.pulse};
    const int ng[3] = {a.npump, a.npot, a.npulse};
    for (int q = 0; q < 3; q++)
        for (int g = 0; g < ng[q]; g++) {
            amp0[q][g] = std::complex<real>(env[q]->temporal(g, t_k));
            amp1[q][g] = std::complex<real>(env[q]->temporal(g, s->rnd(t_k + dt)));
        }
    auto set_amps = [&](real w) {
        for (int q = 0; q < 3; q++)
            for (int g = 0; g < ng[q]; g++) {
                const real re = amp0[q][g].real() * (real(1) - w) + w * amp1[q][g].real();
                const real im = amp0[q][g].imag() * (real(1) - w) + w * amp1[q][g].imag();
                if (q == 0) a.pump_amp[g] = re, a.pump_amp_im[g] = im;
                v2 z;
                z.x = re;
                z.y = im;
                if (q == 1) a.pot_amp[g] = z;
                if (q == 2) a.pulse_amp[g] = z;
            }
    };
    /* weights of a sum: the nonzero a[s][j] (or b, e) with their slots */
    auto set_weights = [&](const double* wts, int n, int* slot, real* w, int& count, real* wsum) {
        count = 0;
        real sum = 0;
        for (int j = 0; j < n; j++) {
            sum += (real)wts[j];
            if (wts[j] == 0.0) continue;
            slot[count] = j;
            w[count] = (real)wts[j];
            count++;
        }
        if (wsum) *wsum = sum;
    };
    const int H = s->staged_H;
    for (int st = 0; st < tab.S; st++) {
        set_amps((real)tab.c[st]);
        a.stage = st;
        a.ext = H - (st + 1);
        PHX_HIP(phx::launch_staged<real>(a, phx::StagedOp::K, s->stream));
        if (st + 1 < tab.S) {
            set_weights(tab.a[st + 1], st + 1, a.slot, a.w, a.nw, &a.wsum);
            PHX_HIP(phx::launch_staged<real>(a, phx::StagedOp::SUM, s->stream));
        }
    }
    set_weights(tab.b, tab.NB, a.slot, a.w, a.nw, &a.wsum);
    a.ne = 0;
    if (tab.err) set_weights(tab.e, tab.S, a.eslot, a.e, a.ne, nullptr);
    PHX_HIP(phx::launch_staged<real>(a, phx::StagedOp::FINAL, s->stream));
    if (tab.err)
        PHX_HIP(phx::launch_reduce_partials(s->st_partials, (int)phx::staged_final_blocks(c.N_c, c.N_r), s->sums, s->stream));
    return PHX_OK;
}

/* one step attempt of the configured iterator: fused generic kernel or staged path */
template <typename real>
int rk_attempt(phx_solver* s, phx::RkTableau tab, double t_k, double dt) {
    return s->staged ? staged_attempt<real>(s, tab, t_k, dt) : generic_launch<real>(s, tab, t_k, dt);
}

/* Solver::applyFFTFilter(true) (solver_fft.cu:67-100) on psi+ and, with -tetm, psi-. */
template <typename real>
int fft_filter_t(phx_solver* s) {
    auto& c = s->cfg;
    if (!s->tw_c || !s->tw_r) return fail(PHX_ERR_STATE, "no FFT mask in this configuration");
    using v2 = typename phx::Vec2<real>::type;
    for (int k = 0; k < (c.tetm ? 2 : 1); k++)
        PHX_HIP(phx::launch_fft_filter<real>((v2*)s->psi[k][s->cur], s->tw_c, s->tw_r, (const real*)s->mask[k], c.N_c, c.N_r,
                                             s->stream, (v2*)s->spec[k]));
    s->spec_raw = false;
    return PHX_OK;
}

/* applyFFTFilter(false) (solver_fft.cu:67-79): without a mask the reference only transforms psi for
   output; here psi is copied and the transform is taken when the spectrum is read (spectrum_ready) */
template <typename real>
int fft_snapshot_t(phx_solver* s) {
    const size_t bytes = 2 * s->cells() * sizeof(real);
    for (int k = 0; k < 2; k++)
        if (s->spec[k]) PHX_HIP(hipMemcpyAsync(s->spec[k], s->psi[k][s->cur], bytes, hipMemcpyDeviceToDevice, s->stream));
    s->spec_raw = true;
    return PHX_OK;
}

template <typename real>
int spectrum_ready_t(phx_solver* s) {
    if (!s->spec_raw) return PHX_OK;
    using v2 = typename phx::Vec2<real>::type;
    for (int k = 0; k < 2; k++)
        if (s->spec[k]) PHX_HIP(phx::launch_fft_forward<real>((v2*)s->spec[k], s->tw_c, s->tw_r, s->cfg.N_c, s->cfg.N_r, s->stream));
    s->spec_raw = false;
    return PHX_OK;
}

/* SSFM::step (split_step_fourier_method.cu:25-71): linear half step in k-space, nonlinear
   full step in real space, linear half step, noise.  The linear factor of a cell is
   exp(-i hbar k^2 / (2 m) dt / 2) / (N_c N_r) with k = 2 * 3.1415926535 * (index or
   index - N) / L (gp_scalar_linear_fourier, kernel_gp_compute.cuh:244-262); it is separable
   and kept as a row table and a column table, remade when dt changes.  psi is updated in
   place; the reservoir goes through the spare buffer. */
template <typename real>
int ssfm_step(phx_solver* s, double t_pre) {
    auto& c = s->cfg;
    using v2 = typename phx::Vec2<real>::type;
    if (s->ssfm_dt != s->dt) {
        const double dt = s->dt, lin = c.h_bar_s / 2.0 / c.m_eff;
        std::vector<std::complex<double>> rowf(c.N_r), colf(c.N_c);
        for (int r = 0; r < c.N_r; r++) {
            const double ky = 2.0 * 3.1415926535 * (double)(r <= c.N_r / 2 ? r : r - c.N_r) / c.L_y;
            rowf[r] = std::polar(1.0, -lin * ky * ky * dt / 2.0) / ((double)c.N_c * (double)c.N_r);
        }
        for (int col = 0; col < c.N_c; col++) {
            const double kx = 2.0 * 3.1415926535 * (double)(col <= c.N_c / 2 ? col : col - c.N_c) / c.L_x;
            colf[col] = std::polar(1.0, -lin * kx * kx * dt / 2.0);
        }
        if (int r = upload_complex<real>(s, s->ssfm_rowf, rowf)) return r;
        if (int r = upload_complex<real>(s, s->ssfm_colf, colf)) return r;
        s->ssfm_dt = s->dt;
    }
    const int ncomp = c.tetm ? 2 : 1;
    phx::SsfmArgs<real> a{};
    const size_t N = s->cells();
    /* sources: the group fields (group g at g N) at each group's amplitude of the step start
       (updateKernelTime, solver_iterator.cu:11-27), or the one summed / single-group field */
    static const double w0[1] = {0.0};
    a.npump = s->st_npump ? s->st_npump : (s->pump[0] ? 1 : 0);
    a.npot = s->st_npot ? s->st_npot : (s->pot[0] ? 1 : 0);
    a.npulse = s->st_npulse;
    for (int g = 0; g < a.npump; g++) {
        if (s->st_npump)
            a.pamp[g] = (real)c.pump.temporal(g, t_pre).real();
        else
            stage_amps<real>(c.pump, t_pre, s->dt, w0, 1, &a.pamp[g]);
        /* a complex reservoir takes Im a_g of the group (of the single group when the fields are summed) */
        a.pamp_im[g] = 0;
        if (c.complex_reservoir && (s->st_npump || (c.pump.groupSize() == 1 && !c.pump.allConstant())))
            a.pamp_im[g] = (real)c.pump.temporal(s->st_npump ? g : 0, t_pre).imag();
    }
    for (int g = 0; g < a.npot; g++) {
        if (s->st_npot) {
            const std::complex<real> z(c.potential.temporal(g, t_pre));
            a.vamp[g].x = z.real();
            a.vamp[g].y = z.imag();
        } else {
            stage_amps<real>(c.potential, t_pre, s->dt, w0, 1, &a.vamp[g].x);
            a.vamp[g].y = 0;
        }
    }
    for (int g = 0; g < a.npulse; g++) {
        const std::complex<real> z(c.pulse.temporal(g, t_pre));
        a.lamp[g].x = z.real();
        a.lamp[g].y = z.imag();
    }
    for (int k = 0; k < ncomp; k++) {
        a.psi[k] = (v2*)s->psi[k][s->cur];
        a.n_cur[k] = c.use_reservoir ? (real*)s->n[k][s->cur] : nullptr;
        a.n_nxt[k] = c.use_reservoir ? (real*)s->n[k][s->cur ^ 1] : nullptr;
        for (int g = 0; g < a.npump; g++) a.pump[k][g] = (const real*)s->pump[k] + (size_t)g * N;
        for (int g = 0; g < a.npot; g++) a.pot[k][g] = (const real*)s->pot[k] + (size_t)g * N;
        for (int g = 0; g < a.npulse; g++) a.pulse[k][g] = (const v2*)s->st_pulse[k] + (size_t)g * N;
    }
    a.noise = (const v2*)s->noise;
    a.cells = N;
    a.tetm = c.tetm;
    a.stochastic = c.use_stochastic;
    a.cres = c.complex_reservoir;
    a.dt = (real)s->dt;
    a.ph = phys_of<real>(s);
    for (int k = 0; k < ncomp; k++)
        PHX_HIP(phx::launch_fft_propagate<real>(a.psi[k], s->tw_c, s->tw_r, s->ssfm_rowf, s->ssfm_colf, c.N_c, c.N_r, s->stream));
    PHX_HIP(phx::launch_ssfm_nonlinear<real>(a, s->stream));
    for (int k = 0; k < ncomp; k++)
        PHX_HIP(phx::launch_fft_propagate<real>(a.psi[k], s->tw_c, s->tw_r, s->ssfm_rowf, s->ssfm_colf, c.N_c, c.N_r, s->stream));
    PHX_HIP(phx::launch_ssfm_independent<real>(a, s->stream));
    return PHX_OK;
}

/* Autotuner of the fused RK4 launch shape: time each candidate tile height on this
   grid (writing into the spare buffers, so the state is untouched) and keep the
   fastest.  Results are cached per (precision, grid, feature set) for the process.
   PHX_RK4_TY fixes the shape; PHX_AUTOTUNE=0 keeps the default. */
template <typename real>
int autotune_rk4(phx_solver* s) {
    static std::map<std::string, phx::Rk4Tune> cache;
    auto& c = s->cfg;
    if (std::getenv("PHX_RK4_TY")) return PHX_OK;
    if (const char* e = std::getenv("PHX_AUTOTUNE"))
        if (std::atoi(e) == 0) return PHX_OK;
    const std::string key = std::to_string(sizeof(real)) + ":" + std::to_string(c.N_c) + "x" + std::to_string(c.N_r) + ":" +
                            std::to_string(c.use_reservoir) + std::to_string(c.use_pumps) + std::to_string(c.use_potentials);
    if (auto it = cache.find(key); it != cache.end()) {
        s->tune = it->second;
        return PHX_OK;
    }
    hipEvent_t e0, e1;
    PHX_HIP(hipEventCreate(&e0));
    PHX_HIP(hipEventCreate(&e1));
    const int cands[] = {8, 10, 12, 16, 20, 24};
    /* deeper row prefetch exists for the pumped reservoir model */
    const int pfs[] = {2, 3, 4, 6, 8};
    const int npf = (c.use_reservoir && c.use_pumps && !c.use_potentials) ? 5 : 1;
    /* three round-robin passes over the candidates, each candidate's fastest pass counts: a
       single pass in a fixed order would time the first candidates while the clocks ramp up */
    std::vector<float> cand_ms(6 * npf, 1e30f);
    for (int pass = 0; pass < 3; pass++) {
        for (int cand = 0; cand < 6 * npf; cand++) {
            phx::Rk4Tune t = s->tune;
            t.ty = cands[cand % 6];
            t.pf = pfs[cand / 6];
            const phx::Rk4Tune keep = s->tune;
            s->tune = t;
            int rc = PHX_OK;
            /* each launch reads the current buffer and writes the other; flipping cur back keeps the state */
            for (int w = 0; w < 2 && !rc; w++) rc = rk4_launch<real>(s, s->t, s->dt), s->cur ^= 1;
            PHX_HIP(hipEventRecord(e0, s->stream));
            for (int r = 0; r < 6 && !rc; r++) rc = rk4_launch<real>(s, s->t, s->dt), s->cur ^= 1;
            PHX_HIP(hipEventRecord(e1, s->stream));
            PHX_HIP(hipEventSynchronize(e1));
            s->tune = keep;
            if (rc) return rc;
            float ms = 0;
            PHX_HIP(hipEventElapsedTime(&ms, e0, e1));
            cand_ms[cand] = std::min(cand_ms[cand], ms);
        }
    }
    int best = 0;
    for (int cand = 1; cand < 6 * npf; cand++)
        if (cand_ms[cand] < cand_ms[best]) best = cand;
    phx::Rk4Tune best_t = s->tune;
    best_t.ty = cands[best % 6];
    best_t.pf = pfs[best / 6];
    PHX_HIP(hipEventDestroy(e0));
    PHX_HIP(hipEventDestroy(e1));
    s->tune = best_t;
    cache[key] = best_t;
    return PHX_OK;
}

/* The step itself: the noise of the step, then Solver::step(variable_time_step) of the
   configured iterator (runge_kutta.cu:74-94, dormand_prince.cu:37-95, ...) at kernel time
   t_pre with s->dt; rejected adaptive attempts are retried at kernel time t_adv (the time
   iterate() has advanced to, which adaptTimeStep's updateKernelTime uploads, solver.cu:97),
   and s->dt leaves as adaptTimeStep sets it.  No t, iteration or FFT-guard bookkeeping. */
template <typename real>
int autotune_flow(phx_solver* s);
int flow_check(phx_solver* s);

template <typename real>
int step_body_t(phx_solver* s, double t_pre, double t_adv, bool force_fixed) {
    auto& c = s->cfg;
    if (c.use_stochastic) {
        /* the noise of this step is drawn once, with sqrt(dt), before any attempt (solver.cu:36-45) */
        using v2 = typename phx::Vec2<real>::type;
        PHX_HIP(phx::launch_noise_field<real>((v2*)s->noise, c.N_c, c.N_r, c.random_seed, (uint64_t)s->iteration,
                                              (real)(c.stochastic_amplitude * std::sqrt(s->dt)), s->stream));
    }
    const bool adaptive = c.adaptive && s->iter && s->iter->can_adapt && !force_fixed;
    const bool imag = c.imag_time_amplitude != 0.0;
    if (s->ssfm) {
        if (int r = ssfm_step<real>(s, t_pre)) return r;
    } else if (c.iterator == "RK4" && !c.tetm && !c.use_stochastic && !imag && !s->staged) {
        if (!s->tuned) {
            if (int r = autotune_rk4<real>(s)) return r;
            if (int r = autotune_flow<real>(s)) return r;
            s->tuned = true;
        }
        if (int r = rk4_launch<real>(s, t_pre, s->dt)) return r;
    } else if (!adaptive) {
        if (int r = rk_attempt<real>(s, s->iter->fixed, t_pre, s->dt)) return r;
        s->cur ^= 1;
    } else {
        /* step(true) of the adaptive iterators (dormand_prince.cu:37-95, bogacki.cu:28-75,
           cash_karp.cu:31-83, fehlberg.cu:33-70,100-152) + Solver::adaptTimeStep (solver.cu:72-105) */
        int rejections = 0;
        for (;;) {
            /* retries see the kernel time already advanced (updateKernelTime in adaptTimeStep, solver.cu:97) */
            const double t_k = rejections == 0 ? t_pre : t_adv;
            if (int r = rk_attempt<real>(s, s->iter->adaptive, t_k, s->dt)) return r;
            PHX_HIP(hipMemcpyAsync(s->sums_host, s->sums, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
            PHX_HIP(hipStreamSynchronize(s->stream));
            double new_dt = 0, final_error = 0;
            const bool accept = adapt_dt(s->sums_host[0], s->sums_host[1], c.tolerance, s->dt, c.dt_min, c.dt_max,
                                         s->iter->power, &new_dt, &final_error);
            s->last_error = final_error;
            s->dt = s->rnd(new_dt);
            if (accept) break;
            s->rejected++;
            if (++rejections > s->max_rejections)
                return fail(PHX_ERR_STEP, "adaptive step rejected " + std::to_string(rejections) + " times (error " +
                                              std::to_string(final_error) + " >= tolerance at dt = " + std::to_string(s->dt) + ")");
        }
        s->cur ^= 1; /* swapBuffers (solver.cu:107-114) */
    }
    return PHX_OK;
}

/* Solver::iterate (solver.cu:22-70).  Returns 1 if a step was taken, 0 at t_max, <0 on error. */
template <typename real>
int iterate_t(phx_solver* s, bool force_fixed) {
    auto& c = s->cfg;
    if (s->t >= c.t_max) return 0;
    const double t_pre = s->t;
    s->t = s->rnd(s->t + s->dt);   /* t += dt before the step (solver.cu:50) */
    if (int r = step_body_t<real>(s, t_pre, s->t, force_fixed)) return r;
    const bool imag = c.imag_time_amplitude != 0.0;
    if (imag) {
        /* normalizeImaginaryTimePropagation after the step (solver.cu:55-56) */
        using v2 = typename phx::Vec2<real>::type;
        PHX_HIP(phx::launch_imag_normalize<real>((v2*)s->psi[0][s->cur], c.tetm ? (v2*)s->psi[1][s->cur] : nullptr,
                                                 s->cells(), c.imag_time_amplitude, c.dV, s->norm_partials, s->norm_sums,
                                                 s->stream));
        /* the reservoir to the same amplitude (solver_normalize_imag_time.cu:16-28).  With the Runge-Kutta
           iterators imaginary time makes it complex (the layout of psi); the split-step one keeps it real,
           normalised here as N_c N_r / 2 pairs of reals (N_c, N_r are even): the same sum of squares and the
           same scale on every element */
        if (c.use_reservoir) {
            const size_t ncell = c.complex_reservoir ? s->cells() : s->cells() / 2;
            PHX_HIP(phx::launch_imag_normalize<real>((v2*)s->n[0][s->cur], c.tetm ? (v2*)s->n[1][s->cur] : nullptr, ncell,
                                                     c.imag_time_amplitude, c.dV, s->norm_partials, s->norm_sums, s->stream));
        }
    }
    s->iteration++;
    if (s->rnd(s->t - s->fft_cached_t) < c.fft_every) return 1; /* FFT guard in the build's real (solver.cu:62-63) */
    s->fft_cached_t = s->t;
    /* without a mask the reference only transforms for output (solver_fft.cu:78-79): no state change */
    if (c.use_fft_mask) {
        if (int r = fft_filter_t<real>(s)) return r;
    } else if (s->spec[0]) {
        if (int r = fft_snapshot_t<real>(s)) return r;
    }
    return 1;
}

int iterate(phx_solver* s, bool force_fixed) {
    return s->precision == PHX_FP32 ? iterate_t<float>(s, force_fixed) : iterate_t<double>(s, force_fixed);
}

constexpr int kGraphSteps = 50; /* graph replay: even, so the buffer parity returns */

/* One launch of rk4_scalar_flow for `steps` (even) steps of the same run; returns 1 when the
   kernel cannot take this grid or launch shape (the caller falls back to the graph). */
constexpr int kFlowSpinLimit = 1 << 21; /* polls of about 1 us: a wait of ~2 s ends the launch */

/* Kernel arguments of a fixed RK4 step of the scalar model from psi/n `in` to `out`, with the
   per-stage temporal amplitudes at the handle's t (constant over a multi-step launch). */
template <typename real>
phx::Rk4Args<real> rk4_args(phx_solver* s, void* psi_in, void* n_in, void* psi_out, void* n_out) {
    auto& c = s->cfg;
    using v2 = typename phx::Vec2<real>::type;
    phx::Rk4Args<real> a{};
    a.psi_in = (const v2*)psi_in;
    a.psi_out = (v2*)psi_out;
    a.n_in = c.use_reservoir ? (const real*)n_in : nullptr;
    a.n_out = c.use_reservoir ? (real*)n_out : nullptr;
    a.pump = (const real*)s->pump[0];
    a.pot = (const real*)s->pot[0];
    a.Nc = c.N_c;
    a.Nr = c.N_r;
    a.periodic_x = c.periodic_x;
    a.periodic_y = c.periodic_y;
    a.alt_rows = phx::rk4_alt_rows();
    a.dt = (real)s->dt;
    static const double w[4] = {0.0, 0.5, 0.5, 1.0};
    stage_amps<real>(c.pump, s->t, s->dt, w, 4, a.pamp);
    stage_amps<real>(c.potential, s->t, s->dt, w, 4, a.vamp_h);
    for (int i = 0; i < 4; i++) a.vamp_h[i] = (real)((double)a.vamp_h[i] * c.one_over_h_bar_s);
    a.ph = phys_of<real>(s);
    return a;
}

/* Counters of the multi-step kernel, allocated on first use once the kernel is known to take
   this grid with the handle's tile shape.  Returns 1 when it does not (or PHX_RK4_FLOW=0). */
template <typename real>
int flow_ready(phx_solver* s) {
    if (s->flow_off) return 1;
    if (const char* e = std::getenv("PHX_RK4_FLOW"))
        if (std::atoi(e) == 0) return 1; /* PHX_RK4_FLOW=0: one launch per step */
    auto& c = s->cfg;
    const size_t counters = (size_t)((c.N_c + 55) / 56) * (size_t)((c.N_r + 7) / 8) * phx::kFlowCounterStride;
    if (!s->flow_counters) {
        const phx::Rk4Args<real> a = rk4_args<real>(s, s->psi[0][s->cur], s->n[0][s->cur], s->psi[0][s->cur ^ 1], s->n[0][s->cur ^ 1]);
        if (phx::launch_rk4_flow<real>(a, s->tune, phx::Rk4Flow{nullptr, nullptr, 0, 2, 0}, s->stream) != hipSuccess) {
            (void)hipGetLastError();
            s->flow_off = true;
            return 1;
        }
        if (int r = dev_alloc(s, (void**)&s->flow_counters, counters * sizeof(int))) return r; /* any tile height >= 8 */
        if (int r = dev_alloc(s, (void**)&s->flow_error, sizeof(int))) return r;
        s->flow_base = 0;
    }
    if (s->flow_base > (1 << 30)) { /* keep the counters far from overflow */
        PHX_HIP(hipMemsetAsync(s->flow_counters, 0, counters * sizeof(int), s->stream));
        s->flow_base = 0;
    }
    return PHX_OK;
}

/* One launch of the multi-step kernel: `steps` steps from psi/n `in` (`out` is the second buffer).
   hipSuccess or the launcher's error (nothing ran). */
template <typename real>
hipError_t flow_launch(phx_solver* s, const phx::Rk4Tune& tune, void* psi_in, void* n_in, void* psi_out, void* n_out, int steps) {
    int spin_limit = kFlowSpinLimit;
    if (const char* e = std::getenv("PHX_FLOW_SPIN_LIMIT")) spin_limit = std::atoi(e); /* tests of the time-out path */
    const phx::Rk4Args<real> a = rk4_args<real>(s, psi_in, n_in, psi_out, n_out);
    const phx::Rk4Flow f{s->flow_counters, s->flow_error, s->flow_base, steps, spin_limit};
    const hipError_t e = phx::launch_rk4_flow<real>(a, tune, f, s->stream);
    if (e == hipSuccess) {
        s->flow_base += steps;
    } else {
        (void)hipGetLastError();
    }
    return e;
}

template <typename real>
int flow_block_t(phx_solver* s, int steps) {
    if (int r = flow_ready<real>(s)) return r;
    if (flow_launch<real>(s, s->tune, s->psi[0][s->cur], s->n[0][s->cur], s->psi[0][s->cur ^ 1], s->n[0][s->cur ^ 1], steps) !=
        hipSuccess) {
        s->flow_off = true; /* nothing ran: step by graph from now on */
        return 1;
    }
    s->flow_steps += steps;
    s->flow_launches++;
    return PHX_OK;
}

/* Tile shape of the multi-step kernel, timed on that kernel: every candidate runs kFlowTuneSteps
   steps from the second buffer through a temporary one and back (an even count), so the state in
   the current buffer is not touched.  Two passes, each candidate's faster one counts; shapes the
   kernel does not take are skipped.  The winner replaces the one-step tuning's choice. */
constexpr int kFlowTuneSteps = 40;

template <typename real>
int autotune_flow(phx_solver* s) {
    static std::map<std::string, phx::Rk4Tune> cache;
    auto& c = s->cfg;
    if (std::getenv("PHX_RK4_TY")) return PHX_OK;
    if (const char* e = std::getenv("PHX_AUTOTUNE"))
        if (std::atoi(e) == 0) return PHX_OK;
    const auto constant = [](const phx::Envelope& e) { return e.groupSize() != 1 || e.allConstant(); };
    if (s->staged || c.iterator != "RK4" || c.tetm || c.use_stochastic || c.imag_time_amplitude != 0.0 || !constant(c.pump) ||
        !constant(c.potential))
        return PHX_OK;
    const std::string key = std::to_string(sizeof(real)) + ":" + std::to_string(c.N_c) + "x" + std::to_string(c.N_r) + ":" +
                            std::to_string(c.use_reservoir) + std::to_string(c.use_pumps) + std::to_string(c.use_potentials) +
                            std::to_string(c.periodic_x) + std::to_string(c.periodic_y);
    if (auto it = cache.find(key); it != cache.end()) {
        s->tune = it->second;
        return PHX_OK;
    }
    const phx::Rk4Tune one_step = s->tune;
    s->tune.ty = 10; /* a shape the kernel has, for the availability check */
    s->tune.pf = 2;
    const int fr = flow_ready<real>(s);
    s->tune = one_step;
    if (fr) return fr < 0 ? fr : PHX_OK;
    const size_t N = s->cells();
    void *tmp_psi = nullptr, *tmp_n = nullptr;
    PHX_HIP(hipMalloc(&tmp_psi, 2 * N * sizeof(real)));
    if (c.use_reservoir && hipMalloc(&tmp_n, N * sizeof(real)) != hipSuccess) {
        (void)hipFree(tmp_psi);
        return fail(PHX_ERR_HIP, "allocation of the tuning buffer failed");
    }
    hipEvent_t e0, e1;
    PHX_HIP(hipEventCreate(&e0));
    PHX_HIP(hipEventCreate(&e1));
    const int tys[] = {8, 10, 12, 16, 20};
    const int pfs[] = {2, 3, 4, 6, 8};
    const int npf = (c.use_reservoir && c.use_pumps && !c.use_potentials) ? 5 : 1;
    std::vector<float> ms(5 * npf, 1e30f);
    void* psi2 = s->psi[0][s->cur ^ 1];
    void* n2 = s->n[0][s->cur ^ 1];
    int rc = PHX_OK;
    for (int pass = 0; pass < 2 && !rc; pass++) {
        for (int cand = 0; cand < 5 * npf && !rc; cand++) {
            phx::Rk4Tune t = one_step;
            t.ty = tys[cand % 5];
            t.pf = pfs[cand / 5];
            if (hipEventRecord(e0, s->stream) != hipSuccess) rc = fail(PHX_ERR_HIP, "hipEventRecord");
            if (!rc && flow_launch<real>(s, t, psi2, n2, tmp_psi, tmp_n, kFlowTuneSteps) != hipSuccess) continue;
            float m = 0;
            if (!rc && (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&m, e0, e1) != hipSuccess))
                rc = fail(PHX_ERR_HIP, "event timing");
            if (!rc) ms[cand] = std::min(ms[cand], m);
        }
    }
    (void)hipStreamSynchronize(s->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(tmp_psi);
    if (tmp_n) (void)hipFree(tmp_n);
    if (rc) return rc;
    if (int r = flow_check(s)) return r;
    int best = -1;
    for (int cand = 0; cand < 5 * npf; cand++)
        if (ms[cand] < 1e29f && (best < 0 || ms[cand] < ms[best])) best = cand;
    if (best < 0) {
        s->flow_off = true;
        return PHX_OK;
    }
    s->tune.ty = tys[best % 5];
    s->tune.pf = pfs[best / 5];
    cache[key] = s->tune;
    return PHX_OK;
}

/* The error word of rk4_scalar_flow, read after the stream has drained.  After a time-out the
   counters no longer match flow_base: the multi-step launch stays off for this handle. */
int flow_check(phx_solver* s) {
    if (!s->flow_error) return PHX_OK;
    int err = 0;
    PHX_HIP(hipMemcpy(&err, s->flow_error, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        s->flow_off = true;
        PHX_HIP(hipMemset(s->flow_error, 0, sizeof(int))); /* reported once */
        return fail(PHX_ERR_HIP, "multi-step RK4 launch: a tile waited too long for its neighbours; the state is undefined");
    }
    return PHX_OK;
}

/* Runs of fixed RK4 steps of the scalar model whose kernel arguments do not change from step
   to step (constant temporal amplitudes, no noise) leave the one-launch-per-step path: up to
   kFlowMaxSteps of them as one launch of the multi-step kernel, or, where that kernel does not
   take the grid, kGraphSteps of them replayed from a captured hipGraph.  A run stops before the
   step that reaches t_max, that fires the k-space filter, or that would pass t_stop (the
   caller shortens that one); the host advances t and the FFT guard exactly as iterate()
   would.  *taken: the steps run (0: not eligible, the caller iterates). */
constexpr int64_t kFlowMaxSteps = 500;

template <typename real>
int block_t(phx_solver* s, int64_t max_steps, double t_stop, int64_t* taken) {
    auto& c = s->cfg;
    *taken = 0;
    const auto constant = [](const phx::Envelope& e) { return e.groupSize() != 1 || e.allConstant(); };
    if (!s->tuned || s->staged || c.iterator != "RK4" || c.tetm || c.use_stochastic || c.imag_time_amplitude != 0.0 ||
        !constant(c.pump) || !constant(c.potential))
        return PHX_OK;
    double t = s->t, fc = s->fft_cached_t, t_graph = 0, fc_graph = 0;
    int64_t m = 0;
    const int64_t lim = std::min(max_steps, kFlowMaxSteps);
    while (m < lim) {
        if (t >= c.t_max || t >= t_stop) break;
        const double tn = s->rnd(t + s->dt);
        if (tn > t_stop) break;
        double fcn = fc;
        if (!(s->rnd(tn - fc) < c.fft_every)) {
            if (c.use_fft_mask || s->spec[0]) break;
            fcn = tn;
        }
        t = tn;
        fc = fcn;
        if (++m == kGraphSteps) t_graph = t, fc_graph = fc;
    }
    if (m >= 2) {
        const int fr = flow_block_t<real>(s, (int)m);
        if (fr < 0) return fr;
        if (fr == PHX_OK) {
            if (m & 1) s->cur ^= 1; /* an odd run leaves the result in the other buffer */
            s->t = t;
            s->fft_cached_t = fc;
            s->iteration += m;
            *taken = m;
            return PHX_OK;
        }
    }
    if (m < kGraphSteps || s->graph_off) return PHX_OK;
    if (const char* e = std::getenv("PHX_GRAPH"))
        if (std::atoi(e) == 0) return PHX_OK; /* PHX_GRAPH=0: launch by launch */
    if (!s->graph || s->graph_parity != s->cur || s->graph_dt != s->dt || s->graph_stream != s->stream) {
        if (s->graph) PHX_HIP(hipGraphExecDestroy(s->graph));
        s->graph = nullptr;
        const int cur = s->cur;
        if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            s->graph_off = true; /* e.g. the legacy default stream: step launch by launch */
            return PHX_OK;
        }
        int rc = PHX_OK;
        for (int k = 0; k < kGraphSteps && !rc; k++) rc = rk4_launch<real>(s, s->t, s->dt);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(s->stream, &g);
        s->cur = cur;
        if (rc) return rc;
        PHX_HIP(ce);
        const hipError_t ie = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PHX_HIP(ie);
        s->graph_parity = cur;
        s->graph_dt = s->dt;
        s->graph_stream = s->stream;
    }
    PHX_HIP(hipGraphLaunch(s->graph, s->stream));
    s->t = t_graph;
    s->fft_cached_t = fc_graph;
    s->iteration += kGraphSteps;
    *taken = kGraphSteps;
    return PHX_OK;
}

int field_info(phx_solver* s, int field, void*** slot, size_t* elems) {
    const size_t N = s->cells();
    const int k = (field & 1);
    if (k == 1 && !s->cfg.tetm && field <= PHX_POTENTIAL_MINUS) return fail(PHX_ERR_INVALID, "minus component requires -tetm");
    static void* none = nullptr;
    switch (field) {
        case PHX_PSI_PLUS:
        case PHX_PSI_MINUS: *slot = &s->psi[k][s->cur]; *elems = 2 * N; break;
        case PHX_N_PLUS:
        case PHX_N_MINUS: *slot = &s->n[k][s->cur]; *elems = s->cfg.complex_reservoir ? 2 * N : N; break;
        case PHX_PUMP_PLUS:
        case PHX_PUMP_MINUS: *slot = &s->pump[k]; *elems = N; break;
        case PHX_POTENTIAL_PLUS:
        case PHX_POTENTIAL_MINUS: *slot = &s->pot[k]; *elems = N; break;
        case PHX_FFT_MASK_PLUS:
        case PHX_FFT_MASK_MINUS: *slot = &s->mask[k]; *elems = N; break;
        case PHX_FFT_PLUS:
        case PHX_FFT_MINUS: *slot = &s->spec[k]; *elems = 2 * N; break;
        default: *slot = &none; return fail(PHX_ERR_INVALID, "unknown field id " + std::to_string(field));
    }
    if (**slot == nullptr) return fail(PHX_ERR_INVALID, "field " + std::to_string(field) + " is not allocated in this configuration");
    return PHX_OK;
}

} // namespace

extern "C" {

const char* phx_last_error(void) { return g_err.c_str(); }

const char* phx_version(void) { return "phoenix_amd 0.1 (gfx950)"; }

int phx_create(phx_solver** out, int precision, int argc, const char* const* argv) {
    if (!out) return fail(PHX_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (precision != PHX_FP32 && precision != PHX_FP64) return fail(PHX_ERR_INVALID, "precision must be PHX_FP32 or PHX_FP64");
    std::vector<std::string> args;
    for (int i = 0; i < argc; i++) args.push_back(argv[i] ? argv[i] : "");
    if (args.empty()) args.push_back("phoenix");
    auto* s = new phx_solver();
    s->precision = precision;
    s->rsize = precision == PHX_FP32 ? 4 : 8;
    std::string e = phx::parse_config(s->cfg, args, precision == PHX_FP32);
    if (e.empty()) e = phx::finalize_config(s->cfg);
    if (!e.empty()) {
        delete s;
        return fail(e.find("not provided") != std::string::npos || e.find("not yet") != std::string::npos ? PHX_ERR_UNSUPPORTED : PHX_ERR_INVALID, e);
    }
    *out = s;
    g_err.clear();
    return PHX_OK;
}

int phx_initialize(phx_solver* s) {
    if (!s) return fail(PHX_ERR_INVALID, "null handle");
    if (s->initialized) return fail(PHX_ERR_STATE, "already initialized");
    if (!s->own_stream) PHX_HIP(hipStreamCreate(&s->own_stream));
    if (!s->stream) s->stream = s->own_stream;
    return s->precision == PHX_FP32 ? initialize_t<float>(s) : initialize_t<double>(s);
}

int phx_step(phx_solver* s, int64_t n_steps, int64_t* n_done) {
    if (!s || !s->initialized) return fail(PHX_ERR_STATE, "phx_step before phx_initialize");
    int64_t done = 0;
    while (done < n_steps) {
        if (n_steps - done >= 2) {
            int64_t taken = 0;
            const double inf = std::numeric_limits<double>::infinity();
            const int r = s->precision == PHX_FP32 ? block_t<float>(s, n_steps - done, inf, &taken)
                                                   : block_t<double>(s, n_steps - done, inf, &taken);
            if (r < 0) {
                if (n_done) *n_done = done;
                return r;
            }
            if (taken) {
                done += taken;
                continue;
            }
        }
        int r = iterate(s, false);
        if (r < 0) {
            if (n_done) *n_done = done;
            return r;
        }
        if (r == 0) break;
        done++;
    }
    if (n_done) *n_done = done;
    return PHX_OK;
}

int phx_step_at(phx_solver* s, double t, double* dt, int variable_time_step, int64_t* rejected) {
    if (!s || !s->initialized) return fail(PHX_ERR_STATE, "phx_step_at before phx_initialize");
    if (!dt || !(*dt > 0.0)) return fail(PHX_ERR_INVALID, "phx_step_at needs dt > 0");
    const int64_t rej0 = s->rejected;
    const double t_keep = s->t;
    s->dt = s->rnd(*dt);
    const double t_pre = s->rnd(t);
    const double t_adv = s->rnd(t_pre + s->dt);
    const int r = s->precision == PHX_FP32 ? step_body_t<float>(s, t_pre, t_adv, variable_time_step == 0)
                                           : step_body_t<double>(s, t_pre, t_adv, variable_time_step == 0);
    s->t = t_keep;
    if (rejected) *rejected = s->rejected - rej0;
    if (r) return r;
    s->iteration++; /* the noise stream's step index */
    *dt = s->dt;
    return PHX_OK;
}

/* main.cu:80-117 (render disabled): iterate in chunks of output_every, shortening the
   step that would overshoot the next output time. */
int phx_run_interval(phx_solver* s, int64_t* n_done, int* finished) {
    if (!s || !s->initialized) return fail(PHX_ERR_STATE, "phx_run_interval before phx_initialize");
    auto& c = s->cfg;
    int64_t done = 0;
    if (s->t < c.t_max) {
        /* main.cu:84-104: iterate up to the next output time; the step that would pass it is
           shortened to land on it and runs with a fixed step.  The reference evaluates the
           output time, t + dt and the shortened dt in its real type; out_iter * output_every
           is exact in double, so rounding it once gives the real-type product. */
        const double target = s->rnd((double)s->out_iter * c.output_every);
        while (s->t < target) {
            int64_t taken = 0;
            const int br = s->precision == PHX_FP32 ? block_t<float>(s, kFlowMaxSteps, target, &taken)
                                                    : block_t<double>(s, kFlowMaxSteps, target, &taken);
            if (br < 0) {
                if (n_done) *n_done = done;
                return br;
            }
            if (taken) {
                done += taken;
                continue;
            }
            const double dt = s->dt;
            bool forced = false;
            if (s->rnd(s->t + s->dt) > target) {
                const double next_dt = target - s->t;
                if (next_dt > 0) {
                    s->dt = s->rnd(next_dt);
                    forced = true;
                }
            }
            int r = iterate(s, forced);
            if (forced) s->dt = dt;
            if (r < 0) {
                if (n_done) *n_done = done;
                return r;
            }
            if (r == 0) break;
            done++;
        }
        s->out_iter++;
    }
    if (n_done) *n_done = done;
    if (finished) *finished = s->t >= c.t_max ? 1 : 0;
    return PHX_OK;
}

int phx_run(phx_solver* s, int64_t* n_done) {
    if (!s || !s->initialized) return fail(PHX_ERR_STATE, "phx_run before phx_initialize");
    int64_t done = 0;
    int finished = 0;
    while (!finished) {
        int64_t d = 0;
        const int r = phx_run_interval(s, &d, &finished);
        done += d;
        if (r != PHX_OK) {
            if (n_done) *n_done = done;
            return r;
        }
    }
    if (n_done) *n_done = done;
    return PHX_OK;
}

int phx_apply_fft_filter(phx_solver* s) {
    if (!s || !s->initialized) return fail(PHX_ERR_STATE, "not initialized");
    return s->precision == PHX_FP32 ? fft_filter_t<float>(s) : fft_filter_t<double>(s);
}

int phx_get_field(phx_solver* s, int field, void* host, size_t bytes) {
    if (!s || !s->initialized) return fail(PHX_ERR_STATE, "not initialized");
    void** slot;
    size_t elems;
    if (int r = field_info(s, field, &slot, &elems)) return r;
    if (bytes != elems * s->rsize) return fail(PHX_ERR_INVALID, "byte count mismatch: expected " + std::to_string(elems * s->rsize));
    if (field == PHX_FFT_PLUS || field == PHX_FFT_MINUS)
        if (int r = s->precision == PHX_FP32 ? spectrum_ready_t<float>(s) : spectrum_ready_t<double>(s)) return r;
    PHX_HIP(hipMemcpyAsync(host, *slot, bytes, hipMemcpyDeviceToHost, s->stream));
    PHX_HIP(hipStreamSynchronize(s->stream));
    return flow_check(s);
}

int phx_set_field(phx_solver* s, int field, const void* host, size_t bytes) {
    if (!s || !s->initialized) return fail(PHX_ERR_STATE, "not initialized");
    if (field == PHX_FFT_PLUS || field == PHX_FFT_MINUS) return fail(PHX_ERR_INVALID, "the k-space spectra are read-only");
    void** slot;
    size_t elems;
    if (int r = field_info(s, field, &slot, &elems)) return r;
    if (bytes != elems * s->rsize) return fail(PHX_ERR_INVALID, "byte count mismatch: expected " + std::to_string(elems * s->rsize));
    PHX_HIP(hipMemcpyAsync(*slot, host, bytes, hipMemcpyHostToDevice, s->stream));
    PHX_HIP(hipStreamSynchronize(s->stream));
    return PHX_OK;
}

int phx_device_ptr(phx_solver* s, int field, void** dptr) {
    if (!s || !s->initialized || !dptr) return fail(PHX_ERR_STATE, "not initialized");
    void** slot;
    size_t elems;
    if (int r = field_info(s, field, &slot, &elems)) return r;
    if (field == PHX_FFT_PLUS || field == PHX_FFT_MINUS)
        if (int r = s->precision == PHX_FP32 ? spectrum_ready_t<float>(s) : spectrum_ready_t<double>(s)) return r;
    *dptr = *slot;
    return PHX_OK;
}

int phx_get_time(phx_solver* s, double* t, double* dt) {
    if (!s) return fail(PHX_ERR_INVALID, "null handle");
    if (t) *t = s->initialized ? s->t : s->cfg.t;
    if (dt) *dt = s->initialized ? s->dt : s->cfg.dt;
    return PHX_OK;
}

int phx_set_time(phx_solver* s, double t, double dt) {
    if (!s) return fail(PHX_ERR_INVALID, "null handle");
    if (dt <= 0) return fail(PHX_ERR_INVALID, "dt must be positive");
    s->t = s->rnd(t);
    s->dt = s->rnd(dt);
    s->cfg.t = s->t;
    s->cfg.dt = s->dt;
    return PHX_OK;
}

static const phx::Envelope* source_envelope(phx_solver* s, int source) {
    switch (source) {
        case PHX_SRC_PUMP: return &s->cfg.pump;
        case PHX_SRC_POTENTIAL: return &s->cfg.potential;
        case PHX_SRC_PULSE: return &s->cfg.pulse;
        case PHX_SRC_FFT_MASK: return &s->cfg.fft_mask;
    }
    return nullptr;
}

int phx_source_groups(phx_solver* s, int source, int* n_groups) {
    if (!s || !n_groups) return fail(PHX_ERR_INVALID, "null argument");
    const phx::Envelope* e = source_envelope(s, source);
    if (!e) return fail(PHX_ERR_INVALID, "unknown source " + std::to_string(source));
    *n_groups = e->size() > 0 ? e->groupSize() : 0;
    return PHX_OK;
}

int phx_source_field(phx_solver* s, int source, int component, int group, double* out, size_t bytes) {
    if (!s || !out) return fail(PHX_ERR_INVALID, "null argument");
    const phx::Envelope* e = source_envelope(s, source);
    if (!e) return fail(PHX_ERR_INVALID, "unknown source " + std::to_string(source));
    const auto& c = s->cfg;
    if (component < 0 || component > 1 || group < 0 || group >= e->groupSize())
        return fail(PHX_ERR_INVALID, "no such component / temporal group");
    if (bytes != 2 * s->cells() * sizeof(double)) return fail(PHX_ERR_INVALID, "buffer size mismatch");
    std::vector<std::complex<double>> f(s->cells());
    const std::string err = e->calculate(f.data(), group, component == 0 ? phx::POL_PLUS : phx::POL_MINUS, c.N_c, c.N_r, c.L_x,
                                         c.L_y, c.dx, c.dy, source == PHX_SRC_FFT_MASK ? 1.0 : 0.0);
    if (!err.empty()) return fail(PHX_ERR_INVALID, err);
    std::memcpy(out, f.data(), bytes);
    return PHX_OK;
}

int phx_source_temporal(phx_solver* s, int source, int group, double t, double* re, double* im, double* t0, double* freq,
                        double* sigma, int* constant) {
    if (!s) return fail(PHX_ERR_INVALID, "null handle");
    const phx::Envelope* e = source_envelope(s, source);
    if (!e) return fail(PHX_ERR_INVALID, "unknown source " + std::to_string(source));
    if (group < 0 || group >= e->groupSize()) return fail(PHX_ERR_INVALID, "no such temporal group");
    const auto& T = e->groups[group];
    const std::complex<double> v = e->temporal(group, t);
    if (re) *re = v.real();
    if (im) *im = v.imag();
    if (t0) *t0 = T.t0;
    if (freq) *freq = T.freq;
    if (sigma) *sigma = T.sigma;
    if (constant) *constant = (T.kind & phx::TMP_CONSTANT) ? 1 : 0;
    return PHX_OK;
}

int phx_get_output_config(phx_solver* s, phx_output_config* oc) {
    if (!s || !oc) return fail(PHX_ERR_INVALID, "null argument");
    const auto& c = s->cfg;
    std::memset(oc, 0, sizeof(*oc));
    std::snprintf(oc->path, sizeof(oc->path), "%s", c.path.c_str());
    std::string keys;
    for (size_t i = 0; i < c.output_keys.size(); i++) keys += (i ? "," : "") + c.output_keys[i];
    std::snprintf(oc->keys, sizeof(oc->keys), "%s", keys.c_str());
    oc->history = c.history;
    oc->hx0 = c.history_x0;
    oc->hx1 = c.history_x1 < 0 ? c.N_c : c.history_x1;
    oc->hy0 = c.history_y0;
    oc->hy1 = c.history_y1 < 0 ? c.N_r : c.history_y1;
    oc->hinc = std::max(1, c.history_increment);
    oc->hevery = std::max(1, c.history_every);
    oc->hstart = c.history_start;
    oc->precision = c.fileout_precision;
    oc->L_x = c.L_x;
    oc->L_y = c.L_y;
    std::snprintf(oc->filetype, sizeof(oc->filetype), "%s", c.output_filetype.c_str());
    std::snprintf(oc->name, sizeof(oc->name), "%s", c.output_name.c_str());
    oc->verbose_io = c.verbose_io;
    return PHX_OK;
}

int phx_get_info(phx_solver* s, phx_info* info) {
    if (!s || !info) return fail(PHX_ERR_INVALID, "null argument");
    const auto& c = s->cfg;
    std::memset(info, 0, sizeof(*info));
    info->N_c = c.N_c;
    info->N_r = c.N_r;
    info->precision = s->precision;
    info->tetm = c.tetm;
    info->use_reservoir = c.use_reservoir;
    info->use_pump = c.use_pumps;
    info->use_potential = c.use_potentials;
    info->use_stochastic = c.use_stochastic;
    info->use_fft_mask = c.use_fft_mask;
    info->adaptive = c.adaptive;
    info->periodic_x = c.periodic_x;
    info->periodic_y = c.periodic_y;
    info->halo = c.halo;
    std::snprintf(info->iterator, sizeof(info->iterator), "%s", c.iterator.c_str());
    info->dx = c.dx;
    info->dy = c.dy;
    info->dt = s->initialized ? s->dt : c.dt;
    info->t = s->initialized ? s->t : c.t;
    info->t_max = c.t_max;
    info->fft_every = c.fft_every;
    info->tolerance = c.tolerance;
    info->dt_min = c.dt_min;
    info->dt_max = c.dt_max;
    info->iteration = s->iteration;
    info->rejected = s->rejected;
    info->tile_rows = s->tuned ? s->tune.ty : 0;
    info->prefetch_rows = s->tuned ? s->tune.pf : 0;
    info->staged = s->staged;
    info->multi_step_steps = s->flow_steps;
    info->multi_step_launches = s->flow_launches;
    info->complex_reservoir = s->cfg.complex_reservoir;
    return PHX_OK;
}

int phx_set_stream(phx_solver* s, void* stream) {
    if (!s) return fail(PHX_ERR_INVALID, "null handle");
    s->stream = stream ? (hipStream_t)stream : s->own_stream;
    return PHX_OK;
}

int phx_sync(phx_solver* s) {
    if (!s) return fail(PHX_ERR_INVALID, "null handle");
    if (s->stream) PHX_HIP(hipStreamSynchronize(s->stream));
    return flow_check(s);
}

void phx_destroy(phx_solver* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->sums_host) (void)hipHostFree(s->sums_host);
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
    delete s;
}

double phx_physics_defaults(phx_physics* ph, int N_c, int N_r, double L_x, double L_y) {
    phx::SystemConfig c;
    c.N_c = N_c;
    c.N_r = N_r;
    c.L_x = L_x;
    c.L_y = L_y;
    c.t_max = 1e30;
    phx::finalize_config(c);
    ph->dx = c.dx;
    ph->dy = c.dy;
    ph->dV = c.dV;
    ph->m_eff_scaled = c.m_eff_scaled;
    ph->one_over_h_bar_s = c.one_over_h_bar_s;
    ph->gamma_c = c.gamma_c;
    ph->gamma_r = c.gamma_r;
    ph->g_c = c.g_c;
    ph->g_r = c.g_r;
    ph->R = c.R;
    ph->g_pm = c.g_pm;
    ph->delta_LT = c.delta_LT;
    return c.magic_timestep;
}

} /* extern "C" */

static phx::SystemConfig cfg_from_physics(const phx_physics* ph) {
    phx::SystemConfig c;
    c.dx = ph->dx;
    c.dy = ph->dy;
    c.dV = ph->dV;
    c.one_over_dx2 = 1.0 / (ph->dx * ph->dx);
    c.one_over_dy2 = 1.0 / (ph->dy * ph->dy);
    c.m2_over_dx2_p_dy2 = -2.0 * (c.one_over_dx2 + c.one_over_dy2);
    c.m_eff_scaled = ph->m_eff_scaled;
    c.one_over_h_bar_s = ph->one_over_h_bar_s;
    c.gamma_c = ph->gamma_c;
    c.gamma_r = ph->gamma_r;
    c.g_c = ph->g_c;
    c.g_r = ph->g_r;
    c.R = ph->R;
    c.g_pm = ph->g_pm;
    c.delta_LT = ph->delta_LT;
    return c;
}

template <typename real>
static int rk4_device_t(const phx_physics* ph, int N_c, int N_r, int px, int py, void* psi_a, void* n_a, void* psi_b, void* n_b,
                        const void* pump, const void* pot, double dt, int64_t n_steps, hipStream_t st) {
    using v2 = typename phx::Vec2<real>::type;
    const phx::SystemConfig c = cfg_from_physics(ph);
    phx::Rk4Args<real> a{};
    a.Nc = N_c;
    a.Nr = N_r;
    a.periodic_x = px;
    a.periodic_y = py;
    a.alt_rows = phx::rk4_alt_rows();
    a.dt = (real)dt;
    for (int i = 0; i < 4; i++) {
        a.pamp[i] = 1;
        a.vamp_h[i] = (real)c.one_over_h_bar_s;
    }
    a.ph = phx::make_phys<real>(c);
    a.pump = (const real*)pump;
    a.pot = (const real*)pot;
    const phx::Rk4Tune tune = rk4_tune_from_env(phx::Rk4Tune{});
    for (int64_t k = 0; k < n_steps; k++) {
        const bool even = (k % 2) == 0;
        a.psi_in = (const v2*)(even ? psi_a : psi_b);
        a.psi_out = (v2*)(even ? psi_b : psi_a);
        a.n_in = n_a ? (const real*)(even ? n_a : n_b) : nullptr;
        a.n_out = n_a ? (real*)(even ? n_b : n_a) : nullptr;
        PHX_HIP(phx::launch_rk4_scalar<real>(a, tune, st));
    }
    if (n_steps % 2 == 1) {
        PHX_HIP(hipMemcpyAsync(psi_a, psi_b, 2 * sizeof(real) * (size_t)N_c * N_r, hipMemcpyDeviceToDevice, st));
        if (n_a) PHX_HIP(hipMemcpyAsync(n_a, n_b, sizeof(real) * (size_t)N_c * N_r, hipMemcpyDeviceToDevice, st));
    }
    return PHX_OK;
}

/* Double buffer of phx_rk4_scalar_device_noalloc: one array per device, loaded with the code
   object and kept until the process ends.  Host code never names it directly (on the host the
   symbol is only a stand-in): it uses the device address that hipGetSymbolAddress returns,
   looked up once per device. */
constexpr size_t kScratchBytes = (size_t)PHX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static int scratch_address(void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = 0;
    PHX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PHX_ERR_UNSUPPORTED, "device index " + std::to_string(dev));
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(PHX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[dev] = p;
    }
    *out = addr[dev];
    return PHX_OK;
}

extern "C" int phx_rk4_scalar_device_noalloc(int precision, const phx_physics* ph, int N_c, int N_r, int periodic_x,
                                             int periodic_y, void* psi, void* n, const void* pump, const void* pot, double dt,
                                             int64_t n_steps, void* stream) {
    if (!ph || !psi || N_c <= 0 || N_r <= 0 || n_steps < 0) return fail(PHX_ERR_INVALID, "invalid argument");
    if (pump && !n) return fail(PHX_ERR_INVALID, "a pump needs a reservoir");
    if (precision != PHX_FP32 && precision != PHX_FP64) return fail(PHX_ERR_INVALID, "precision");
    const size_t rsize = precision == PHX_FP32 ? 4 : 8, cells = (size_t)N_c * (size_t)N_r;
    const size_t psi_bytes = (2 * cells * rsize + 255) & ~(size_t)255;
    const size_t need = psi_bytes + (n ? cells * rsize : 0);
    if (need > kScratchBytes)
        return fail(PHX_ERR_INVALID, "grid of " + std::to_string(cells) + " cells needs " + std::to_string(need) +
                                         " bytes of scratch; the allocation-free entry point holds " + std::to_string(kScratchBytes));
    void* base = nullptr;
    if (int r = scratch_address(&base)) return r;
    void* psi_b = base;
    void* n_b = n ? (char*)base + psi_bytes : nullptr;
    hipStream_t st = (hipStream_t)stream;
    if (precision == PHX_FP32) return rk4_device_t<float>(ph, N_c, N_r, periodic_x, periodic_y, psi, n, psi_b, n_b, pump, pot, dt, n_steps, st);
    return rk4_device_t<double>(ph, N_c, N_r, periodic_x, periodic_y, psi, n, psi_b, n_b, pump, pot, dt, n_steps, st);
}

extern "C" int phx_rk4_scalar_device(int precision, const phx_physics* ph, int N_c, int N_r, int periodic_x, int periodic_y, void* psi_a,
                          void* n_a, void* psi_b, void* n_b, const void* pump, const void* pot, double dt, int64_t n_steps,
                          void* stream) {
    if (!ph || !psi_a || !psi_b || N_c <= 0 || N_r <= 0 || n_steps < 0) return fail(PHX_ERR_INVALID, "invalid argument");
    if ((n_a == nullptr) != (n_b == nullptr)) return fail(PHX_ERR_INVALID, "n_a and n_b must both be given or both be NULL");
    if (pump && !n_a) return fail(PHX_ERR_INVALID, "a pump needs a reservoir");
    hipStream_t st = (hipStream_t)stream;
    if (precision == PHX_FP32) return rk4_device_t<float>(ph, N_c, N_r, periodic_x, periodic_y, psi_a, n_a, psi_b, n_b, pump, pot, dt, n_steps, st);
    if (precision == PHX_FP64) return rk4_device_t<double>(ph, N_c, N_r, periodic_x, periodic_y, psi_a, n_a, psi_b, n_b, pump, pot, dt, n_steps, st);
    return fail(PHX_ERR_INVALID, "precision");
}

