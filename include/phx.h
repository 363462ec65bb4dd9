/*
 * phx.h -- C ABI of the MI355X-native PHOENIX time-stepping library (libphx.so).
 *
 * The reference exposes this path through three surfaces, none of them a C ABI:
 * the PHOENIX_{cpu,gpu}_{fp32,fp64} executables and their flags
 * (source/system/system_initialization.cpp:13-261, source/system/config_readin.cpp:7-42),
 * the internal C++ Solver interface (include/solver/solver.hpp:141-160 with
 * SolverFactory::create, include/solver/solver_factory.hpp:53-84), and pyphoenix,
 * which spawns the executable (pyphoenix/pyphoenix.py:162-215).  Each entry point
 * below names the reference interface it replaces.
 *
 * Conventions
 *   - Status: 0 = PHX_OK, negative = error; phx_last_error() (thread local) holds
 *     the message.  The library never calls exit(); the CLI maps errors to exit(1)
 *     (system_validate_input.cpp:53-56).
 *   - The solver owns its device memory; callers own host buffers.
 *   - A handle is not thread safe.  Each handle owns one HIP stream (or uses the
 *     one given by phx_set_stream).
 *   - Host field layout: full grid, row-major N_r x N_c, index row*N_c + col.
 *     Complex fields (psi, pulse, k-space spectra) are interleaved {re, im}; real fields (n, P, V,
 *     FFT mask) are plain.  Element type is float for PHX_FP32, double for PHX_FP64.
 */
#ifndef PHX_H
#define PHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phx_solver phx_solver;

enum phx_precision { PHX_FP32 = 0, PHX_FP64 = 1 };

enum phx_status {
    PHX_OK = 0,
    PHX_ERR_INVALID = -1,      /* invalid input (system_validate_input.cpp:19-56)      */
    PHX_ERR_HIP = -2,          /* HIP runtime failure                                   */
    PHX_ERR_UNSUPPORTED = -3,  /* flag / iterator outside this library's path           */
    PHX_ERR_IO = -4,           /* file could not be read                                */
    PHX_ERR_STEP = -5,         /* adaptive step rejected more than the allowed count    */
    PHX_ERR_STATE = -6         /* call out of order (e.g. step before initialize)       */
};

enum phx_field {
    PHX_PSI_PLUS = 0,
    PHX_PSI_MINUS = 1,
    PHX_N_PLUS = 2,
    PHX_N_MINUS = 3,
    PHX_PUMP_PLUS = 4,
    PHX_PUMP_MINUS = 5,
    PHX_POTENTIAL_PLUS = 6,
    PHX_POTENTIAL_MINUS = 7,
    PHX_FFT_MASK_PLUS = 8,
    PHX_FFT_MASK_MINUS = 9,
    /* k-space spectra (complex, read-only; matrix.fft_plus / fft_minus, solver_fft.cu:67-100): after a
       filter application DFT(psi) / (N_c N_r) * fftshift(mask), without a mask the unnormalised DFT(psi)
       at the last FFT time; kept when fft_every < t_max and the output keys ask for fft output
       (solver_output_matrices.cu:41,62), row-major with k = 0 first, zero before the first FFT time */
    PHX_FFT_PLUS = 10,
    PHX_FFT_MINUS = 11
};

/* Summary of a configured solver (read-only view of SystemParameters,
   include/system/system_parameters.hpp:25-64). */
typedef struct {
    int32_t N_c, N_r;
    int32_t precision;          /* phx_precision */
    int32_t tetm;               /* -tetm */
    int32_t use_reservoir, use_pump, use_potential, use_stochastic, use_fft_mask;
    int32_t adaptive;           /* -adaptive with an adaptive iterator */
    int32_t periodic_x, periodic_y;
    int32_t halo;               /* stages of the step (= halo in cells): 4 (RK4), 6 (DP45), 7 (DP45 -adaptive) */
    char iterator[16];          /* registered --iterator key: "RK4", "DP45", "Heun2", "BoSha", ... */
    double dx, dy, dt, t, t_max, fft_every, tolerance, dt_min, dt_max;
    int64_t iteration;          /* accepted steps so far */
    int64_t rejected;           /* rejected adaptive attempts so far */
    int32_t tile_rows;          /* output rows per wavefront tile of the fused RK4 kernel (0 before the first step) */
    int32_t prefetch_rows;      /* rows it loads ahead of use */
    int32_t staged;             /* 1: the stage-by-stage path runs the step (K's in HBM: pulses, several
                                   non-constant temporal groups, complex potential envelopes, DP85, NSRK78) */
    int64_t multi_step_steps;   /* fixed RK4 steps taken by the multi-step kernel (phx_step and
                                   phx_run_interval run constant-source stretches that way, up to 500 steps
                                   per launch; PHX_RK4_FLOW=0: never) */
    int64_t multi_step_launches; /* its launches */
    int32_t complex_reservoir;  /* 1: the reservoir fields are complex ({re, im} per cell: PHX_N_PLUS /
                                   PHX_N_MINUS take 2 N_c N_r reals), as the reference's n becomes under
                                   --imagTime with a reservoir or a complex temporal pump envelope */
} phx_info;

/* Parse PHOENIX flags (argv[0] is the program name and is skipped), including
   --config files, derive constants and validate.  No device work.
   Replaces: SystemParameters(argc, argv) (system.cpp:162-184) +
   readConfigFromFile (config_readin.cpp:7-42) + SolverFactory::create
   (solver_factory.hpp:53-60; unknown --iterator -> PHX_ERR_UNSUPPORTED). */
int phx_create(phx_solver** out, int precision, int argc, const char* const* argv);

/* Allocate device fields, evaluate envelopes (pump, potential, initial state,
   initial reservoir, FFT mask) and upload them.
   Replaces: Solver::initialize / initializeMatricesFromSystem
   (solver_initialization.cu:13-174). */
int phx_initialize(phx_solver* s);

/* Run up to n_steps iterate() calls; stops early when t >= t_max.
   *n_done receives the number of accepted steps taken.  Stretches of fixed RK4 steps of the
   scalar model with constant sources (no filter application or t_max inside) go to the device
   as one launch of up to 500 steps in which every tile waits for its neighbours (bitwise the
   same result as step by step); should
   such a wait ever exceed ~2 s, the next phx_sync / phx_get_field returns PHX_ERR_HIP and the
   state is undefined.  PHX_RK4_FLOW=0 launches step by step instead.
   Replaces: Solver::iterate (solver.cu:22-70) in the main loop (main.cu:80-117). */
int phx_step(phx_solver* s, int64_t n_steps, int64_t* n_done);

/* Iterate until t >= t_max.  Replaces the main loop of main.cu:80-117. */
int phx_run(phx_solver* s, int64_t* n_done);

/* One output interval of the main loop: iterate until t reaches the next multiple of
   --outEvery (the step that would pass it is shortened to land on it) or t_max; *finished
   is set to 1 once t >= t_max.  phx_run is a loop over this call.
   Replaces: the inner while loop of main() (main.cu:84-104). */
int phx_run_interval(phx_solver* s, int64_t* n_done, int* finished);

/* Exactly one step of the configured iterator, without iterate()'s bookkeeping: the noise of
   the step (drawn with sqrt(*dt)), then the stages at kernel time t (the pre-step time that
   updateKernelTime uploads before iterate() advances t) and step *dt; with
   variable_time_step != 0 and an adaptive iterator, rejected attempts are retried at kernel
   time t + *dt with the step size adaptTimeStep chose, until one is accepted.  On return *dt
   holds the step size for the next step (unchanged for fixed steps) and *rejected the
   rejected attempts.  The handle's t, FFT guard and imaginary-time normalisation are not
   touched (the caller's iterate() does those); its step counter advances by one.
   Replaces: the virtual Solver::step(bool variable_time_step) (solver.hpp:141), i.e.
   RungeKutta4::step (runge_kutta.cu:74-94), DormandPrince45::step (dormand_prince.cu:37-95)
   and the other REGISTER_SOLVER iterators, for a Solver subclass that forwards to this library. */
int phx_step_at(phx_solver* s, double t, double* dt, int variable_time_step, int64_t* rejected);

/* Envelope sources of a configuration. */
enum phx_source { PHX_SRC_PUMP = 0, PHX_SRC_POTENTIAL = 1, PHX_SRC_PULSE = 2, PHX_SRC_FFT_MASK = 3 };

/* Number of temporal groups of a source (0 if the source is unset).
   Replaces: Envelope::groupSize (envelope.hpp:101). */
int phx_source_groups(phx_solver* s, int source, int* n_groups);

/* Spatial field of one temporal group of a source for component 0 (plus) or 1 (minus),
   evaluated on the host exactly as at initialisation: full grid, row-major, complex
   interleaved doubles (2 N_c N_r).  Replaces: Envelope::calculate for one group
   (envelope.cpp:332-412), i.e. the matrices outputInitialMatrices writes
   (solver_output_matrices.cu:72-150). */
int phx_source_field(phx_solver* s, int source, int component, int group, double* out, size_t bytes);

/* Temporal envelope of a group at time t and its oscillator parameters (0 for a constant
   group, *constant = 1).  Replaces: Envelope::temporal / updateTemporal (envelope.cpp:466-499)
   as cached by Solver::cacheValues (solver_cache.cu:13-35). */
int phx_source_temporal(phx_solver* s, int source, int group, double t, double* re, double* im, double* t0,
                        double* freq, double* sigma, int* constant);

/* File output settings: --path, --output keys (comma separated; default "mat,scalar",
   system.cpp:65), --historyMatrix x0 x1 y0 y1 increment, --historyTime start every,
   --fileoutPrecision (system_initialization.cpp:53-63,160-181), --outputFileType
   (filehandler.cpp:49-69). */
typedef struct {
    char path[512];
    char keys[256];
    int32_t history;                  /* --historyMatrix given */
    int32_t hx0, hx1, hy0, hy1, hinc; /* region [x0, x1) x [y0, y1), every hinc-th cell */
    int32_t hevery;                   /* every hevery-th output interval */
    double hstart;                    /* from this time on */
    int32_t precision;                /* significant digits, -1: stream default */
    double L_x, L_y;
    char filetype[8];                 /* --outputFileType: "txt", "npy", "ash" or "mat" */
    char name[128];                   /* --name: every output file is <path><name>_<file> (empty: <path><file>) */
    int32_t verbose_io;               /* -verboseIO: report each written matrix */
} phx_output_config;
int phx_get_output_config(phx_solver* s, phx_output_config* oc);

/* Apply the k-space filter now (forward FFT, mask/N, inverse FFT on psi).
   Replaces: Solver::applyFFTFilter(true) (solver_fft.cu:67-100). */
int phx_apply_fft_filter(phx_solver* s);

/* Copy a field to / from host memory (full grid, layout above).  bytes must equal
   N_c*N_r*elements*sizeof(real), elements = 2 for psi, the spectra and (with
   phx_info.complex_reservoir) the reservoir, 1 otherwise.  Replaces the host mirror of CUDAMatrix
   (cuda_matrix.cuh:281-306,375-434) used by the reference's output path. */
int phx_get_field(phx_solver* s, int field, void* host, size_t bytes);
int phx_set_field(phx_solver* s, int field, const void* host, size_t bytes);

/* Device pointer of a field (current buffer; valid until the next step), for
   zero-copy interop (torch tensors over the same HIP device).  No reference
   counterpart: the reference keeps device pointers internal (getDevicePtr,
   cuda_matrix.cuh:445-453). */
int phx_device_ptr(phx_solver* s, int field, void** dptr);

/* Time bookkeeping: t and dt of SystemParameters::p (system_parameters.hpp:33). */
int phx_get_time(phx_solver* s, double* t, double* dt);
int phx_set_time(phx_solver* s, double t, double dt);

int phx_get_info(phx_solver* s, phx_info* info);

/* Use an external HIP stream (hipStream_t passed as void*) for all work of this
   handle; NULL restores the handle's own stream (so the null stream itself cannot be set: a
   caller whose work runs on it orders a stream of its own with events).  The same holds for
   phx_ensemble_set_stream and phx_adjoint_set_stream. */
int phx_set_stream(phx_solver* s, void* stream);

/* Wait for all device work of the handle. */
int phx_sync(phx_solver* s);

void phx_destroy(phx_solver* s);

/* Thread-local message of the last failing call ("" if none). */
const char* phx_last_error(void);

/* Library version string. */
const char* phx_version(void);

/* ---------------------------------------------------------------------------
 * Device-pointer entry points (no handle).  Fields are device pointers in the
 * layout above; the caller owns them.  Used by the batched-ensemble and adjoint
 * forms and by the benchmark.  `stream` is a hipStream_t (NULL = default).
 * ------------------------------------------------------------------------- */

/* Physical constants for the device entry points, as derived by
   SystemParameters::calculateAuto (system.cpp:116-160). */
typedef struct {
    double dx, dy, dV;
    double m_eff_scaled, one_over_h_bar_s;
    double gamma_c, gamma_r, g_c, g_r, R, g_pm, delta_LT;
} phx_physics;

/* Fill phx_physics from the reference defaults (system.cpp:18-72) for a grid
   of N_c x N_r cells over L_x x L_y micrometres; returns the magic dt. */
double phx_physics_defaults(phx_physics* ph, int N_c, int N_r, double L_x, double L_y);

/* `n_steps` fixed RK4 steps of the scalar GP + reservoir model on one grid.
   psi_a/n_a hold the state and receive the result; psi_b/n_b are scratch of
   the same size (double buffering).  pump / pot may be NULL.  Replaces one
   RungeKutta4::step per step (runge_kutta.cu:74-94). */
int phx_rk4_scalar_device(int precision, const phx_physics* ph, int N_c, int N_r, int periodic_x, int periodic_y,
                          void* psi_a, void* n_a, void* psi_b, void* n_b, const void* pump, const void* pot,
                          double dt, int64_t n_steps, void* stream);

/* Scratch held by phx_rk4_scalar_device_noalloc on each device (bytes). */
#define PHX_NOALLOC_SCRATCH_BYTES (32u << 20)

/* phx_rk4_scalar_device without the caller's second buffer and without any device allocation:
   psi / n hold the state and receive the result; the double buffer is a fixed array of
   PHX_NOALLOC_SCRATCH_BYTES in the library's device image (each device that runs it holds it
   until the process ends).  Grids needing more than that (2 N_c N_r reals for psi, rounded up
   to 256 bytes, plus N_c N_r for n) are rejected with PHX_ERR_INVALID: up to 2.79 M cells in
   FP32 with a reservoir (e.g. 1024 x 1024, 1600 x 1600), 1.39 M in FP64 (1024 x 1024).  Calls no
   hipMalloc / hipFree; an odd step count copies the result back with hipMemcpyAsync on `stream`.
   All calls on one device share that device's array: calls that overlap in time on one device
   (other streams, other threads) race.  Results are bitwise those of phx_rk4_scalar_device.
   No counterpart in the reference: for callers that cannot allocate device memory (code under
   hipGraph capture, tight serving loops, harnesses that forbid hipMalloc). */
int phx_rk4_scalar_device_noalloc(int precision, const phx_physics* ph, int N_c, int N_r, int periodic_x, int periodic_y,
                                  void* psi, void* n, const void* pump, const void* pot, double dt, int64_t n_steps,
                                  void* stream);

/* ---------------------------------------------------------------------------
 * Batched ensembles: many independent scalar grids of different sizes stepped
 * by fixed RK4 in one launch per step.  Each instance is the system that one
 * reference process integrates with RungeKutta4::step (runge_kutta.cu:74-94,
 * scalar path only); the reference runs one instance per process, so this
 * has no reference entry point of its own.  Every instance shares the physics
 * constants (except dx, dy), the boundary flags, dt and the feature flags, and
 * its result is bitwise that of stepping it alone with phx_rk4_scalar_device.
 * ------------------------------------------------------------------------- */
typedef struct phx_ensemble phx_ensemble;

enum { PHX_ENS_RESERVOIR = 1, PHX_ENS_PUMP = 2, PHX_ENS_POTENTIAL = 4 };

/* n instances of N_c[i] x N_r[i] cells with grid spacings dx[i], dy[i]
   (NULL: ph->dx / ph->dy for all).  flags: PHX_ENS_* (a pump needs the
   reservoir).  All fields start at zero. */
int phx_ensemble_create(phx_ensemble** out, int precision, const phx_physics* ph, int n, const int* N_c, const int* N_r,
                        const double* dx, const double* dy, int periodic_x, int periodic_y, int flags);

/* Copy one field of instance `inst` (PHX_PSI_PLUS, PHX_N_PLUS, PHX_PUMP_PLUS,
   PHX_POTENTIAL_PLUS) from / to host memory; bytes must equal the field size. */
int phx_ensemble_set_field(phx_ensemble* e, int inst, int field, const void* host, size_t bytes);
int phx_ensemble_get_field(phx_ensemble* e, int inst, int field, void* host, size_t bytes);

/* Device address of the current buffer of a field (changes parity with each
   odd step count). */
int phx_ensemble_device_ptr(phx_ensemble* e, int inst, int field, void** dptr);

/* Output rows per wavefront tile of the batched kernel: 8, 10, 12, 16 or 20.
   Without this call the first step times every height and keeps the fastest. */
int phx_ensemble_set_tile_rows(phx_ensemble* e, int ty);

/* n_steps RK4 steps of every instance.  graph_steps >= 2 captures that many
   (rounded down to even) launches in a hipGraph once and replays it; 0 or 1
   launches step by step.  Asynchronous on the ensemble's stream. */
int phx_ensemble_step(phx_ensemble* e, double dt, int64_t n_steps, int graph_steps);

/* Run on a caller-owned hipStream_t (NULL: the ensemble's own stream).  On a stream that
   cannot be captured the steps are launched one by one instead of replayed from a hipGraph. */
int phx_ensemble_set_stream(phx_ensemble* e, void* stream);
int phx_ensemble_sync(phx_ensemble* e);
void phx_ensemble_destroy(phx_ensemble* e);

/* ---------------------------------------------------------------------------
 * Differentiation: n_steps fixed RK4 steps of `batch` scalar grids of one size
 * forward with a checkpoint every `checkpoint_every` steps, and the discrete
 * adjoint backward (stage states recomputed from the checkpoints), giving the
 * gradient of any objective of the final state with respect to the pump P,
 * the potential V and the initial state.  The forward map is RungeKutta4::step
 * (runge_kutta.cu:74-94) with time-independent pump and potential; the
 * reference itself has no adjoint.  All fields are device pointers, batch
 * major: psi (batch x N_r x N_c, interleaved complex), n (batch x N_r x N_c).
 * P, V: one grid shared by the batch (shared_params != 0; their gradients are
 * summed over the batch) or one grid per item.  Gradients of complex fields
 * are dJ/dRe + i dJ/dIm.  Asynchronous on the handle's stream.
 * ------------------------------------------------------------------------- */
typedef struct phx_adjoint phx_adjoint;

/* flags: PHX_ENS_RESERVOIR / PHX_ENS_PUMP / PHX_ENS_POTENTIAL.  Allocates
   ceil(n_steps / checkpoint_every) + checkpoint_every + 1 batched states. */
int phx_adjoint_create(phx_adjoint** out, int precision, const phx_physics* ph, int N_c, int N_r, int batch, int periodic_x,
                       int periodic_y, int flags, int shared_params, double dt, int64_t n_steps, int checkpoint_every);

/* Run forward from (psi0, n0) to (psi_T, n_T), keeping checkpoints.  pump and
   pot must stay valid and unchanged until the last phx_adjoint_backward. */
int phx_adjoint_forward(phx_adjoint* a, const void* psi0, const void* n0, const void* pump, const void* pot, void* psi_T,
                        void* n_T);

/* Given dJ/d(psi_T) and dJ/d(n_T) (NULL: zero), write dJ/dP, dJ/dV, dJ/d(psi0),
   dJ/d(n0) (any output may be NULL).  Needs a preceding phx_adjoint_forward. */
int phx_adjoint_backward(phx_adjoint* a, const void* g_psi_T, const void* g_n_T, void* g_pump, void* g_pot, void* g_psi0,
                         void* g_n0);

/* Run on a caller-owned hipStream_t (NULL: the handle's own stream).  The backward's recompute
   stream waits for this stream's earlier work. */
int phx_adjoint_set_stream(phx_adjoint* a, void* stream);
int phx_adjoint_sync(phx_adjoint* a);
void phx_adjoint_destroy(phx_adjoint* a);

#ifdef __cplusplus
}
#endif
#endif /* PHX_H */
