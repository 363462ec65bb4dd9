/*
 * phx_cpu_rk4.c -- timing restatement of the reference's CPU RK4 step for the scalar GP +
 * reservoir + pump model with periodic boundaries, in float or double (REAL, set at compile
 * time: libphx_cpu_rk4_fp32.so / libphx_cpu_rk4_fp64.so).
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py reports next to the GPU figure,
 * checked against the FP64 oracle (tests/test_oracle_cpu_rk4.py).  The product never loads it.
 *
 * It keeps the structure of the reference's USE_CPU build so that its timing is the reference's
 * CPU cost: the grid split into padded subgrids with a halo of 4 (cuda_matrix.cuh:193-254),
 * one halo synchronisation of psi and n per step (kernel_halo.cuh:50-76, periodic), then one
 * OpenMP parallel loop over subgrids running the whole step sequence of a subgrid
 * (SOLVER_SEQUENCE, cuda_macro.cuh:410-440): K1 on extent 3, the intermediate sum, K2 on 2, sum,
 * K3 on 1, sum, K4 on 0 and the final sum over the padded subgrid (runge_kutta.cu:74-94), each a
 * row loop with an `omp simd` column loop (CALL_SUBGRID_KERNEL, cuda_macro.cuh:363-374).  The
 * arithmetic per cell is that of the oracle's calc_k / sum_to_buffer / final_sum in the same
 * order (kernel_gp_compute.cuh:50-110, kernel_summation.cuh:195-258).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL real;

typedef struct {
    int N, sg, h, ro, N2h, S, sgs; /* grid, subgrid edge, halo, row offset, padded cells, subgrids, per side */
    real *psi, *bpsi, *kpsi;      /* interleaved complex, per subgrid N2h cells; k: 4 slots */
    real *n, *bn, *kn, *pump;
    /* physics */
    real m_eff_scaled, ohb, m2, idx2, idy2, g_c, g_r, R, gamma_c, gamma_r, dt;
} cpu_ws;

static cpu_ws W;

/* padded-subgrid index of the interior cell (r, c) of the full grid */
static inline size_t sub_index(const cpu_ws* w, int r, int c) {
    const int g = (r / w->sg) * w->sgs + c / w->sg;
    return (size_t)g * w->N2h + (size_t)(w->h + r % w->sg) * w->ro + (w->h + c % w->sg);
}

/* halo of every subgrid from the periodic neighbours (kernel_halo.cuh:50-76), psi and n */
static void sync_halos(cpu_ws* w) {
    const int h = w->h, sg = w->sg, N = w->N;
#pragma omp parallel for schedule(static)
    for (int g = 0; g < w->S; g++) {
        const int gr = g / w->sgs, gc = g % w->sgs;
        for (int tr = 0; tr < sg + 2 * h; tr++)
            for (int tc = 0; tc < sg + 2 * h; tc++) {
                if (tr >= h && tr < sg + h && tc >= h && tc < sg + h) continue;
                const int r = (gr * sg + tr - h + N) % N, c = (gc * sg + tc - h + N) % N;
                const size_t dst = (size_t)g * w->N2h + (size_t)tr * w->ro + tc, src = sub_index(w, r, c);
                w->psi[2 * dst] = w->psi[2 * src];
                w->psi[2 * dst + 1] = w->psi[2 * src + 1];
                w->n[dst] = w->n[src];
            }
    }
}

/* K_index on extent halo - index, input psi/n (from_buffer = 0) or the buffer (1) */
static inline void calc_k(const cpu_ws* w, size_t base, int index, int from_buffer) {
    const int ch = w->h - index, ro = w->ro;
    const real* in = from_buffer ? w->bpsi : w->psi;
    const real* inn = from_buffer ? w->bn : w->n;
    real* kp = w->kpsi + 2 * (size_t)(index - 1) * w->S * w->N2h;
    real* kn = w->kn + (size_t)(index - 1) * w->S * w->N2h;
    for (int r = w->h - ch; r < w->h + w->sg + ch; r++) {
        const size_t rb = base + (size_t)r * ro;
#pragma omp simd
        for (int c = w->h - ch; c < w->h + w->sg + ch; c++) {
            const size_t i = rb + c;
            const real pr = in[2 * i], pi = in[2 * i + 1];
            const real mr = pi, mi = -pr; /* -i psi */
            const real nrm = pr * pr + pi * pi;
            const real hr = w->m_eff_scaled * (w->m2 * pr + (in[2 * (i + ro)] + in[2 * (i - ro)]) * w->idy2 +
                                               (in[2 * (i + 1)] + in[2 * (i - 1)]) * w->idx2);
            const real hi = w->m_eff_scaled * (w->m2 * pi + (in[2 * (i + ro) + 1] + in[2 * (i - ro) + 1]) * w->idy2 +
                                               (in[2 * (i + 1) + 1] + in[2 * (i - 1) + 1]) * w->idx2);
            real wr = hi * w->ohb, wi = -hr * w->ohb;
            wr += w->ohb * w->g_c * nrm * mr;
            wi += w->ohb * w->g_c * nrm * mi;
            wr -= (real)0.5 * w->gamma_c * pr;
            wi -= (real)0.5 * w->gamma_c * pi;
            const real nv = inn[i];
            wr += w->ohb * w->g_r * nv * mr;
            wi += w->ohb * w->g_r * nv * mi;
            wr += (real)0.5 * w->R * nv * pr;
            wi += (real)0.5 * w->R * nv * pi;
            real k = -(w->gamma_r + w->R * nrm) * nv;
            k += w->pump[i];
            kn[i] = k;
            kp[2 * i] = wr;
            kp[2 * i + 1] = wi;
        }
    }
}

/* buffer = X + dt * weight * K_index on extent halo - index */
static inline void sum_to_buffer(cpu_ws* w, size_t base, int index, real weight) {
    const int ch = w->h - index, ro = w->ro;
    const real* kp = w->kpsi + 2 * (size_t)(index - 1) * w->S * w->N2h;
    const real* kn = w->kn + (size_t)(index - 1) * w->S * w->N2h;
    const real dt = w->dt;
    for (int r = w->h - ch; r < w->h + w->sg + ch; r++) {
        const size_t rb = base + (size_t)r * ro;
#pragma omp simd
        for (int c = w->h - ch; c < w->h + w->sg + ch; c++) {
            const size_t i = rb + c;
            w->bpsi[2 * i] = w->psi[2 * i] + dt * (weight * kp[2 * i]);
            w->bpsi[2 * i + 1] = w->psi[2 * i + 1] + dt * (weight * kp[2 * i + 1]);
            w->bn[i] = w->n[i] + dt * (weight * kn[i]);
        }
    }
}

/* X += dt * sum_j b_j K_j over the padded subgrid */
static inline void final_sum(cpu_ws* w, size_t base) {
    const real b0 = (real)(1.0 / 6.0), b1 = (real)(1.0 / 3.0), b2 = (real)(1.0 / 3.0), b3 = (real)(1.0 / 6.0);
    const size_t slot = (size_t)w->S * w->N2h;
    const real dt = w->dt;
#pragma omp simd
    for (size_t i = base; i < base + (size_t)w->N2h; i++) {
        real s0 = b0 * w->kpsi[2 * i], s1 = b0 * w->kpsi[2 * i + 1], sn = b0 * w->kn[i];
        s0 += b1 * w->kpsi[2 * (i + slot)];
        s1 += b1 * w->kpsi[2 * (i + slot) + 1];
        sn += b1 * w->kn[i + slot];
        s0 += b2 * w->kpsi[2 * (i + 2 * slot)];
        s1 += b2 * w->kpsi[2 * (i + 2 * slot) + 1];
        sn += b2 * w->kn[i + 2 * slot];
        s0 += b3 * w->kpsi[2 * (i + 3 * slot)];
        s1 += b3 * w->kpsi[2 * (i + 3 * slot) + 1];
        sn += b3 * w->kn[i + 3 * slot];
        w->psi[2 * i] += dt * s0;
        w->psi[2 * i + 1] += dt * s1;
        w->n[i] += dt * sn;
    }
}

static void rk4_step(cpu_ws* w) {
    sync_halos(w);
#pragma omp parallel for schedule(static)
    for (int g = 0; g < w->S; g++) {
        const size_t base = (size_t)g * w->N2h;
        calc_k(w, base, 1, 0);
        sum_to_buffer(w, base, 1, (real)0.5);
        calc_k(w, base, 2, 1);
        sum_to_buffer(w, base, 2, (real)0.5);
        calc_k(w, base, 3, 1);
        sum_to_buffer(w, base, 3, (real)1.0);
        calc_k(w, base, 4, 1);
        final_sum(w, base);
    }
}

static void* xalloc(size_t bytes) {
    void* p = NULL;
    if (posix_memalign(&p, 64, bytes ? bytes : 64)) return NULL;
    memset(p, 0, bytes);
    return p;
}

void cpu_rk4_free(void) {
    free(W.psi); free(W.bpsi); free(W.kpsi); free(W.n); free(W.bn); free(W.kn); free(W.pump);
    memset(&W, 0, sizeof(W));
}

/* Set up an N x N periodic grid split into subgrids_per_side^2 padded subgrids; fields are
   full-grid row-major doubles (psi interleaved complex), converted to REAL.  Constants as the
   oracle's orc_params (system.cpp:116-160).  Returns 0, or -1 for an invalid split. */
int cpu_rk4_setup(int N, int subgrids_per_side, const double* psi, const double* n, const double* pump, double dt,
                  double m_eff_scaled, double one_over_h_bar_s, double m2_over_dx2_p_dy2, double one_over_dx2,
                  double one_over_dy2, double g_c, double g_r, double R, double gamma_c, double gamma_r) {
    cpu_rk4_free();
    if (subgrids_per_side <= 0 || N % subgrids_per_side || N / subgrids_per_side < 4) return -1;
    W.N = N;
    W.sgs = subgrids_per_side;
    W.sg = N / subgrids_per_side;
    W.h = 4;
    W.ro = W.sg + 2 * W.h;
    W.N2h = W.ro * W.ro;
    W.S = subgrids_per_side * subgrids_per_side;
    const size_t cells = (size_t)W.S * W.N2h;
    W.psi = xalloc(2 * cells * sizeof(real));
    W.bpsi = xalloc(2 * cells * sizeof(real));
    W.kpsi = xalloc(8 * cells * sizeof(real));
    W.n = xalloc(cells * sizeof(real));
    W.bn = xalloc(cells * sizeof(real));
    W.kn = xalloc(4 * cells * sizeof(real));
    W.pump = xalloc(cells * sizeof(real));
    if (!W.psi || !W.bpsi || !W.kpsi || !W.n || !W.bn || !W.kn || !W.pump) return -2;
    W.m_eff_scaled = (real)m_eff_scaled;
    W.ohb = (real)one_over_h_bar_s;
    W.m2 = (real)m2_over_dx2_p_dy2;
    W.idx2 = (real)one_over_dx2;
    W.idy2 = (real)one_over_dy2;
    W.g_c = (real)g_c;
    W.g_r = (real)g_r;
    W.R = (real)R;
    W.gamma_c = (real)gamma_c;
    W.gamma_r = (real)gamma_r;
    W.dt = (real)dt;
    for (int r = 0; r < N; r++)
        for (int c = 0; c < N; c++) {
            const size_t i = sub_index(&W, r, c), f = (size_t)r * N + c;
            W.psi[2 * i] = (real)psi[2 * f];
            W.psi[2 * i + 1] = (real)psi[2 * f + 1];
            W.n[i] = (real)n[f];
            W.pump[i] = (real)pump[f];
        }
    /* the pump's halo is constant: synchronised once (initializeMatricesFromSystem) */
    for (int g = 0; g < W.S; g++) {
        const int gr = g / W.sgs, gc = g % W.sgs;
        for (int tr = 0; tr < W.ro; tr++)
            for (int tc = 0; tc < W.ro; tc++) {
                const int r = (gr * W.sg + tr - W.h + N) % N, c = (gc * W.sg + tc - W.h + N) % N;
                W.pump[(size_t)g * W.N2h + (size_t)tr * W.ro + tc] = W.pump[sub_index(&W, r, c)];
            }
    }
    return 0;
}

/* `steps` RK4 steps; *us_each (may be NULL) receives the wall time of each step in microseconds. */
int cpu_rk4_run(int steps, double* us_each) {
    if (!W.psi) return -1;
    for (int s = 0; s < steps; s++) {
        const double t0 = omp_get_wtime();
        rk4_step(&W);
        if (us_each) us_each[s] = 1e6 * (omp_get_wtime() - t0);
    }
    return 0;
}

/* Full-grid state out (doubles). */
int cpu_rk4_get(double* psi, double* n) {
    if (!W.psi) return -1;
    for (int r = 0; r < W.N; r++)
        for (int c = 0; c < W.N; c++) {
            const size_t i = sub_index(&W, r, c), f = (size_t)r * W.N + c;
            psi[2 * f] = (double)W.psi[2 * i];
            psi[2 * f + 1] = (double)W.psi[2 * i + 1];
            n[f] = (double)W.n[i];
        }
    return 0;
}

int cpu_rk4_threads(void) { return omp_get_max_threads(); }
