/*
 * phx_oracle.h -- CPU restatement of PHOENIX's explicit Runge-Kutta time step
 * (driven-dissipative Gross-Pitaevskii + reservoir), FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in phoenix_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Layout conventions at this interface (identical to the product C-ABI):
 *   full grid, row-major, N_r rows x N_c columns, index = row * N_c + col;
 *   complex fields interleaved {re, im} (psi), real fields plain (n, P, V).
 *
 * Internally the oracle follows the reference structure literally: the grid
 * is split into subgrids_columns x subgrids_rows padded subgrids with a halo
 * of width k_max, halos are synchronised once per step (periodic copy or
 * zero), and stage s is evaluated on the extent (halo - s)
 * (reference: include/cuda/cuda_macro.cuh:43-49,110-264,410-440).
 */
#ifndef PHX_ORACLE_H
#define PHX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Threads of the OpenMP subgrid loops; n <= 0 only queries.  Returns the count in use. */
int orc_set_threads(int n);

#define ORC_MAX_GROUPS 8

typedef struct {
    /* geometry */
    int N_c, N_r;
    int subgrids_columns, subgrids_rows;
    int periodic_x, periodic_y;
    double dx, dy, dV;
    /* derived constants (reference: source/system/system.cpp:116-160) */
    double one_over_dx2, one_over_dy2, m2_over_dx2_p_dy2;
    double m_eff_scaled, one_over_h_bar_s, h_bar_s, m_eff;
    /* model parameters (reference: include/system/system_parameters.hpp:25-64) */
    double gamma_c, gamma_r, g_c, g_r, R, g_pm, delta_LT;
    /* feature switches (reference: gp_scalar template flags) */
    int tetm, use_reservoir, use_pump, use_potential, use_stochastic;
    double stochastic_amplitude;
    /* temporal amplitudes of pump / potential at t and t+dt
       (reference: kernel_gp_compute.cuh:79,93 amp*(1-w) + w*amp_next) */
    double pump_amp, pump_amp_next, pot_amp, pot_amp_next;
    /* noise stream (own Philox4x32-10 design, see orc_noise_normal2) */
    uint64_t noise_seed;
    /* --imagTime amplitude; != 0: psi advances by -i dt and is normalised after every step
       (cuda_macro.cuh:178,225; solver_normalize_imag_time.cu:13-45) */
    double imag_time_amplitude;
    /* domain lengths (set by orc_derive; the split-step iterator's wave numbers use them) */
    double L_x, L_y;
    /* Source chains (kernel_gp_compute.cuh:68-100,146-175): when src_groups[q] > 0, source q
       (0 pump, 1 potential, 2 pulse) is sum_g field_g * a_g(w) over its temporal groups, with
       field_g = src_field[q][g][component] (pump, potential: N reals; pulse: N complex) and
       a_g(w) = src_amp[q][g] (1 - w) + w src_amp_next[q][g] (complex {re, im}); the pump
       takes the real part (a real reservoir).  The pump and potential chains replace the
       single fields and amplitudes above; the pulse adds (1/hbar) sum_g field_g a_g(w) to k_psi. */
    int src_groups[3];
    const double* src_field[3][ORC_MAX_GROUPS][2];
    double src_amp[3][ORC_MAX_GROUPS][2], src_amp_next[3][ORC_MAX_GROUPS][2];
    /* Complex reservoir (the reference stores n as Type::complex, cuda_matrix.cuh): != 0 makes the
       n arrays of every call complex ({re, im} per cell).  n then takes the pump chain's full complex
       amplitude sum_g P_g a_g(w) (kernel_gp_compute.cuh:70-75), advances by -i dt in imaginary time
       like psi (runge_sum_to_input_k with complex_dt, cuda_macro.cuh:160-178), enters k_psi as
       complex products, sets the noise amplitude through a complex square root
       (kernel_summation.cuh:209-225), and is normalised in imaginary time
       (solver_normalize_imag_time.cu:13-45).  Imaginary time with a reservoir requires it. */
    int complex_reservoir;
} orc_params;

/* Fill derived constants exactly as SystemParameters::calculateAuto does.
   h_bar_s / m_eff < 0 mean "derive from h_bar, e_e, m_e".  Returns the
   magic time step 0.5*dx*dy/(m_e/e*1e8/m_eff).  (system.cpp:116-160) */
double orc_derive(orc_params* p, double L_x, double L_y, double h_bar, double e_e, double m_e, double h_bar_s, double m_eff);

/* Default physical parameters (system.cpp:18-72). */
void orc_defaults(orc_params* p);

/* Fixed-step RK4: `steps` full RungeKutta4::step() calls on the state.
   psi_plus/psi_minus: 2*N doubles; n_plus/n_minus: N doubles (NULL without reservoir);
   pump_*, pot_*: N doubles or NULL.  t0 is the time at the first step start
   (only used by the noise stream).  Returns 0 on success.
   (reference: runge_kutta.cu:74-94, solver.cu:22-70) */
int orc_rk4_run(const orc_params* p, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                double dt, int steps, int64_t step0);

/* One Dormand-Prince 5(4) attempt with fixed dt (dormand_prince.cu:37-95, adaptive branch).
   On return the state holds the 5th-order solution (what swapBuffers would install),
   *err_sum = sum over interior cells of |dt*sum_j (b_j-b*_j) k_j|^2 (psi+ error counted
   twice with -tetm, as ERROR_K reuses the psi+ slots for psi-: cuda_macro.cuh:244),
   *norm = sum over interior cells of |psi+_new|^2.  Both accumulated in FP64. */
int orc_dp45_attempt(const orc_params* p, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                     const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                     double dt, double noise_dt, int64_t step_index, double* err_sum, double* norm);

/* Step-size controller of Solver::adaptTimeStep (solver.cu:72-105) on given sums.
   Returns accept flag, writes the new dt. */
int orc_adapt_dt(double err_sum, double norm, double tolerance, double dt, double dt_min, double dt_max, double* new_dt, double* final_error);

/* Adaptive DP45 driver: iterate() until t >= t_max (solver.cu:22-70 + main.cu loop).
   Returns number of accepted steps or a negative error. *t and *dt are in/out. */
int64_t orc_dp45_run(const orc_params* p, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                     const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                     double* t, double* dt, double t_max, double tolerance, double dt_min, double dt_max,
                     int max_rejections, int64_t* n_rejected);

/* Generic iterators: the step() sequences of the reference iterators
   (solver/iterator/<name>.cu) restated as data and run on the subgridded workspace with
   the reference's halo_size_.  key: a REGISTER_SOLVER key ("Newton", "MP", "Heun2",
   "Ralston", "Heun3", "Ralston3", "RK3", "SSPRK3", "HW", "RK4", "Ralston4", "Rule38",
   "Nystroem", "Fehlberg2", "BoSha", "CashKarp", "Fehlberg5", "DP45"); -3 if unknown.
   orc_iter_info: halo, adaptTimeStep power (0: fixed-step only) and stage counts.
   orc_iter_run: `steps` fixed steps (step(false)).
   orc_iter_attempt: one step(true) attempt, as orc_dp45_attempt (-2 if fixed-step only).
   orc_iter_adaptive_run: as orc_dp45_run with the iterator's power. */
int orc_iter_info(const char* key, int* halo, double* power, int* stages_fixed, int* stages_adaptive);
int orc_iter_run(const orc_params* p, const char* key, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                 const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                 double dt, int steps, int64_t step0);
int orc_iter_attempt(const orc_params* p, const char* key, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                     const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                     double dt, double noise_dt, int64_t step_index, double* err_sum, double* norm);
/* Split-step Fourier iterator: `steps` SSFM::step() calls on full-grid fields
   (split_step_fourier_method.cu:25-71, gp_scalar_linear_fourier / nonlinear / independent,
   kernel_gp_compute.cuh:244-440).  Transforms: unnormalised 2D DFT of the N_r x N_c
   row-major grid (orc_fft2d).  Pump / potential amplitudes pump_amp / pot_amp (the value at
   the step start); source chains (src_groups > 0) replace the single fields, each group at
   src_amp (the amplitude of the step start): potential V_g a_g (complex a_g), pump P_g Re a_g,
   pulse (dt/hbar) F_g a_g in the scalar model and (-i dt/hbar) F_g a_g with -tetm.  With -tetm
   the reservoirs stay at their input values, as in the reference (its independent step never
   installs the new TE-TM reservoir). */
int orc_ssfm_run(const orc_params* p, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                 const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                 double dt, int steps, int64_t step0);

/* Imaginary-time normalisation of full-grid fields: psi_k *= sqrt(amp / (S_k dV)), S_k = sum of
   |psi_k|^2 over the grid (S_k < 1e-10 -> 1); psi_minus may be NULL.
   (solver_normalize_imag_time.cu:13-45) */
void orc_imag_normalize(const orc_params* p, double* psi_plus, double* psi_minus);
/* The same normalisation of the (complex) reservoir fields n_plus / n_minus (NULL: skipped). */
void orc_imag_normalize_reservoir(const orc_params* p, double* n_plus, double* n_minus);
int orc_adapt_dt_pow(double err_sum, double norm, double tolerance, double dt, double dt_min, double dt_max, double power,
                     double* new_dt, double* final_error);
int64_t orc_iter_adaptive_run(const orc_params* p, const char* key, double* psi_plus, double* psi_minus, double* n_plus,
                              double* n_minus, const double* pump_plus, const double* pump_minus, const double* pot_plus,
                              const double* pot_minus, double* t, double* dt, double t_max, double tolerance, double dt_min,
                              double dt_max, int max_rejections, int64_t* n_rejected);

/* k-space filter (solver_fft.cu:67-100, kernel_fft.cuh:15-39):
   psi <- IDFT( DFT(psi) / (N_c*N_r) * fftshift(mask) ), unnormalised transforms,
   forward sign e^{-2 pi i ...}.  mask is the unshifted N_r x N_c real mask. */
int orc_fft_filter(double* psi, const double* mask, int N_c, int N_r);

/* Unnormalised 2D DFT in place; sign = -1 forward, +1 inverse. */
int orc_fft2d(double* data, int N_c, int N_r, int sign);

/* Envelope spatial shape (envelope.cpp:332-412) for one generated component.
   Flag bits are the reference enums (envelope.hpp:37-61): type 1 gauss, 2 outerExponent,
   4 ring, 8 noDivide, 16 local; behaviour 1 add, 2 multiply, 4 replace, 8 adaptive,
   16 complex.  out is complex interleaved and updated in place. */
void orc_envelope_add(double* out, int N_c, int N_r, double L_x, double L_y, double dx, double dy,
                      double amp, double width_x, double width_y, double pos_x, double pos_y,
                      double exponent, int type_flags, double charge_m, double k0_x, double k0_y, int behaviour);

/* Noise: complex standard normal pair for (seed, step, stage-independent cell index)
   via Philox4x32-10 (Salmon et al., SC'11) + Box-Muller.  Same stream as the HIP kernels. */
void orc_noise_normal2(uint64_t seed, uint64_t step, uint32_t cell, int component, double* re, double* im);
void orc_philox4x32_10(uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Initial random state: 0.01*(N(0,1)+iN(0,1)) per cell from the Philox stream,
   row-major (the seeded complex Gaussian noise of the benchmark workloads). */
void orc_gaussian_field(double* psi, int N, double amplitude, uint64_t seed);

/* Discrete adjoint of `steps` fixed RK4 steps of the scalar model (no TE-TM, no
   noise; pump / potential amplitudes from p as in orc_rk4_run).  The forward
   map is that of orc_rk4_run, evaluated on a padded grid; psi_T / n_T (may be
   NULL) receive its final state.  On entry lam_psi (2N, d/dRe + i d/dIm) and
   lam_n (N, may be NULL) hold the cotangent of the final state; on return the
   cotangent of the initial state.  g_pump / g_pot (N each, may be NULL)
   accumulate the cotangents of P and V.  Returns 0, or -2 for unsupported
   settings.  Not a reference function: the check for the HIP adjoint. */
int orc_rk4_adjoint(const orc_params* p, const double* psi0, const double* n0, const double* pump, const double* pot, double dt,
                    int steps, double* lam_psi, double* lam_n, double* g_pump, double* g_pot, double* psi_T, double* n_T);

#ifdef __cplusplus
}
#endif
#endif
