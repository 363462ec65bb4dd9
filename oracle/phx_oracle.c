/*
 * phx_oracle.c -- FP64 CPU restatement of PHOENIX's RK time step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker).  Never linked into or called by
 * the product library.  Build: oracle/Makefile (gcc -std=c11 -O2
 * -ffp-contract=off, as the reference CPU build rounds without contraction).
 *
 * Every function cites the reference unit it restates.  The structure is the
 * reference CPU path: padded subgrids with halo = k_max, one halo sync per
 * step, stage s on extent (halo - s), full-region final sum.
 *
 * Threads: like the reference CPU build, which runs each step's per-subgrid
 * work in an OpenMP loop over subgrids (cuda_macro.cuh:435), the per-subgrid
 * loops here are OpenMP loops over subgrids.  Every subgrid writes only its
 * own padded cells, so results do not depend on the thread count; the error
 * and norm reductions stay serial (fixed order).
 */
#include "phx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* parameters                                                               */
/* ------------------------------------------------------------------------ */

/* Thread count of the OpenMP subgrid loops (1 without OpenMP); n <= 0 keeps the current one. */
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* SystemParameters::SystemParameters() defaults (system.cpp:18-72). */
void orc_defaults(orc_params* p) {
    memset(p, 0, sizeof(*p));
    p->gamma_c = 0.15;
    p->gamma_r = 1.5 * 0.15;
    p->g_c = 3.e-6;
    p->g_r = 2. * 3.e-6;
    p->R = 0.01;
    p->g_pm = -3.e-6 / 5;
    p->delta_LT = 0.025e-3;
    p->N_c = p->N_r = 400;
    p->subgrids_columns = p->subgrids_rows = 1;
    p->imag_time_amplitude = 0.0;
    p->pump_amp = p->pump_amp_next = 1.0;
    p->pot_amp = p->pot_amp_next = 1.0;
}

/* SystemParameters::calculateAuto (system.cpp:116-160). */
double orc_derive(orc_params* p, double L_x, double L_y, double h_bar, double e_e, double m_e, double h_bar_s, double m_eff) {
    if (h_bar_s < 0) h_bar_s = h_bar / e_e * 1e12;
    if (m_eff < 0) m_eff = m_e / e_e * 1e8;
    const double dt_scaling_factor = m_e / e_e * 1e8 / m_eff;
    p->dx = L_x / (p->N_c - 1);
    p->dy = L_y / (p->N_r - 1);
    p->dV = p->dx * p->dy;
    p->one_over_dx2 = 1.0 / (p->dx * p->dx);
    p->one_over_dy2 = 1.0 / (p->dy * p->dy);
    p->m2_over_dx2_p_dy2 = -2.0 * (p->one_over_dx2 + p->one_over_dy2);
    p->h_bar_s = h_bar_s;
    p->m_eff = m_eff;
    p->L_x = L_x;
    p->L_y = L_y;
    p->m_eff_scaled = -0.5 * h_bar_s * h_bar_s / m_eff;
    p->one_over_h_bar_s = 1.0 / h_bar_s;
    return 0.5 * p->dx * p->dy / dt_scaling_factor;
}

/* ------------------------------------------------------------------------ */
/* noise: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) + Box-Muller    */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(uint32_t ctr[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void box_muller(uint32_t a, uint32_t b, double* z0, double* z1) {
    const double u1 = ((double)a + 0.5) * (1.0 / 4294967296.0);
    const double u2 = ((double)b + 0.5) * (1.0 / 4294967296.0);
    const double r = sqrt(-2.0 * log(u1));
    const double ang = 6.283185307179586476925286766559 * u2;
    *z0 = r * cos(ang);
    *z1 = r * sin(ang);
}

/* counter = {cell, step lo, step hi, component}, key = seed (64 bit). */
void orc_noise_normal2(uint64_t seed, uint64_t step, uint32_t cell, int component, double* re, double* im) {
    uint32_t ctr[4] = {cell, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)component};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    box_muller(out[0], out[1], re, im);
}

void orc_gaussian_field(double* psi, int N, double amplitude, uint64_t seed) {
    for (int i = 0; i < N; i++) {
        double re, im;
        orc_noise_normal2(seed, 0xFFFFFFFFull, (uint32_t)i, 0x7fffffff, &re, &im);
        psi[2 * i] = amplitude * re;
        psi[2 * i + 1] = amplitude * im;
    }
}

/* ------------------------------------------------------------------------ */
/* subgridded workspace (cuda_matrix.cuh:193-254, matrix_container.cuh)     */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_params* p;
    int h, sgc, sgr, ro, N2h, S, kmax, ncomp;
    int nel;                                /* doubles per reservoir cell: 1, or 2 (complex reservoir) */
    double *psi[2], *bpsi[2], *kpsi[2];     /* complex: 2 doubles per cell */
    double *n[2], *bn[2], *kn[2];           /* real, or complex with complex_reservoir */
    double *pump[2], *pot[2];               /* real */
    double *xi;                             /* complex noise per padded cell, per subgrid */
    double *src[3][ORC_MAX_GROUPS][2];      /* source chains per group and component (padded) */
    double dt;
} ws_t;

static void* xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

static int ws_init(ws_t* w, const orc_params* p, int kmax) {
    memset(w, 0, sizeof(*w));
    w->p = p;
    w->h = kmax;
    w->kmax = kmax;
    if (p->N_c % p->subgrids_columns || p->N_r % p->subgrids_rows) return -1;
    w->sgc = p->N_c / p->subgrids_columns;
    w->sgr = p->N_r / p->subgrids_rows;
    /* the halo map copies h rows/columns of the neighbour's interior: needs subgrid edge >= halo */
    if (w->sgc < kmax || w->sgr < kmax) return -1;
    w->ro = w->sgc + 2 * w->h;
    w->N2h = w->ro * (w->sgr + 2 * w->h);
    w->S = p->subgrids_columns * p->subgrids_rows;
    w->ncomp = p->tetm ? 2 : 1;
    /* imaginary time advances the reservoir by -i dt: it needs the complex reservoir */
    if (p->imag_time_amplitude != 0.0 && p->use_reservoir && !p->complex_reservoir) return -1;
    w->nel = p->complex_reservoir ? 2 : 1;
    size_t cells = (size_t)w->S * w->N2h;
    for (int c = 0; c < w->ncomp; c++) {
        w->psi[c] = xcalloc(2 * cells, sizeof(double));
        w->bpsi[c] = xcalloc(2 * cells, sizeof(double));
        w->kpsi[c] = xcalloc(2 * cells * kmax, sizeof(double));
        w->n[c] = xcalloc(w->nel * cells, sizeof(double));
        w->bn[c] = xcalloc(w->nel * cells, sizeof(double));
        w->kn[c] = xcalloc(w->nel * cells * kmax, sizeof(double));
        w->pump[c] = xcalloc(cells, sizeof(double));
        w->pot[c] = xcalloc(cells, sizeof(double));
    }
    w->xi = xcalloc(2 * cells, sizeof(double));
    for (int q = 0; q < 3; q++) {
        if (p->src_groups[q] < 0 || p->src_groups[q] > ORC_MAX_GROUPS) return -1;
        for (int g = 0; g < p->src_groups[q]; g++)
            for (int c = 0; c < w->ncomp; c++) w->src[q][g][c] = xcalloc((q == 2 ? 2 : 1) * cells, sizeof(double));
    }
    return 0;
}

static void ws_free(ws_t* w) {
    for (int c = 0; c < w->ncomp; c++) {
        free(w->psi[c]); free(w->bpsi[c]); free(w->kpsi[c]);
        free(w->n[c]); free(w->bn[c]); free(w->kn[c]);
        free(w->pump[c]); free(w->pot[c]);
    }
    free(w->xi);
    for (int q = 0; q < 3; q++)
        for (int g = 0; g < ORC_MAX_GROUPS; g++)
            for (int c = 0; c < 2; c++) free(w->src[q][g][c]);
}

/* full grid -> padded subgrid interiors (kernel_halo.cuh:7-24) */
static void to_sub(const ws_t* w, double* sub, const double* full, int elem) {
    const orc_params* p = w->p;
    for (int r = 0; r < p->N_r; r++)
        for (int c = 0; c < p->N_c; c++) {
            int g = (r / w->sgr) * p->subgrids_columns + c / w->sgc;
            size_t li = (size_t)g * w->N2h + (size_t)(w->h + r % w->sgr) * w->ro + (w->h + c % w->sgc);
            for (int e = 0; e < elem; e++) sub[li * elem + e] = full[((size_t)r * p->N_c + c) * elem + e];
        }
}

/* padded subgrid interiors -> full grid (kernel_halo.cuh:26-43) */
static void to_full(const ws_t* w, const double* sub, double* full, int elem) {
    const orc_params* p = w->p;
    for (int r = 0; r < p->N_r; r++)
        for (int c = 0; c < p->N_c; c++) {
            int g = (r / w->sgr) * p->subgrids_columns + c / w->sgc;
            size_t li = (size_t)g * w->N2h + (size_t)(w->h + r % w->sgr) * w->ro + (w->h + c % w->sgc);
            for (int e = 0; e < elem; e++) full[((size_t)r * p->N_c + c) * elem + e] = sub[li * elem + e];
        }
}

/* Global cell behind a padded position; returns -1 outside a zero-BC domain. */
static long global_cell(const ws_t* w, int g, int tr, int tc) {
    const orc_params* p = w->p;
    int R = g / p->subgrids_columns, C = g % p->subgrids_columns;
    long row = (long)R * w->sgr + tr - w->h;
    long col = (long)C * w->sgc + tc - w->h;
    if (row < 0 || row >= p->N_r) {
        if (!p->periodic_y) return -1;
        row = (row % p->N_r + p->N_r) % p->N_r;
    }
    if (col < 0 || col >= p->N_c) {
        if (!p->periodic_x) return -1;
        col = (col % p->N_c + p->N_c) % p->N_c;
    }
    return row * p->N_c + col;
}

/* Halo synchronisation (kernel_halo.cuh:50-76 with the map of
   solver_initialization.cu:181-220): every halo cell copies the interior cell
   of the neighbouring subgrid in direction (dr,dc); wraps when periodic,
   otherwise writes zero.  Neighbour indices wrap modulo the subgrid count. */
static void sync_halos(const ws_t* w, double* f, int elem) {
    const orc_params* p = w->p;
    const int h = w->h, sgc = w->sgc, sgr = w->sgr;
    #pragma omp parallel for schedule(static)
    for (int g = 0; g < w->S; g++) {
        int R = g / p->subgrids_columns, C = g % p->subgrids_columns;
        for (int tr = 0; tr < sgr + 2 * h; tr++)
            for (int tc = 0; tc < sgc + 2 * h; tc++) {
                int dr = tr < h ? -1 : (tr >= h + sgr ? 1 : 0);
                int dc = tc < h ? -1 : (tc >= h + sgc ? 1 : 0);
                if (!dr && !dc) continue;
                size_t to = (size_t)g * w->N2h + (size_t)tr * w->ro + tc;
                int Rn = R + dr, Cn = C + dc;
                if ((!p->periodic_x && (Cn < 0 || Cn >= p->subgrids_columns)) || (!p->periodic_y && (Rn < 0 || Rn >= p->subgrids_rows))) {
                    for (int e = 0; e < elem; e++) f[to * elem + e] = 0.0;
                    continue;
                }
                Rn = (Rn + p->subgrids_rows) % p->subgrids_rows;
                Cn = (Cn + p->subgrids_columns) % p->subgrids_columns;
                int gn = Rn * p->subgrids_columns + Cn;
                size_t from = (size_t)gn * w->N2h + (size_t)(tr - dr * sgr) * w->ro + (tc - dc * sgc);
                for (int e = 0; e < elem; e++) f[to * elem + e] = f[from * elem + e];
            }
    }
}

/* ------------------------------------------------------------------------ */
/* RHS: Kernel::Compute::gp_scalar (kernel_gp_compute.cuh:18-236)           */
/* ------------------------------------------------------------------------ */
/* -i*z (typedef.cuh:164-166) */
#define MI_RE(re, im) (im)
#define MI_IM(re, im) (-(re))

/* a_g(w) = amp (1 - w) + w amp_next of group g of source q (kernel_gp_compute.cuh:79) */
static void src_amp(const orc_params* p, int q, int g, double wt, double* re, double* im) {
    *re = p->src_amp[q][g][0] * (1.0 - wt) + wt * p->src_amp_next[q][g][0];
    *im = p->src_amp[q][g][1] * (1.0 - wt) + wt * p->src_amp_next[q][g][1];
}

/* potential chain: wf += (1/hbar) (V_g a_g(w)) (-i psi) for every group, in group order
   (kernel_gp_compute.cuh:88-93,146-150); (mr, mi) = -i psi */
static void add_pot_chain(const ws_t* w, int comp, size_t i, double wt, double mr, double mi, double* wr, double* wi) {
    const orc_params* p = w->p;
    for (int g = 0; g < p->src_groups[1]; g++) {
        double ar, ai;
        src_amp(p, 1, g, wt, &ar, &ai);
        const double v = w->src[1][g][comp][i];
        const double tr = p->one_over_h_bar_s * (v * ar), ti = p->one_over_h_bar_s * (v * ai);
        *wr += tr * mr - ti * mi;
        *wi += tr * mi + ti * mr;
    }
}

/* pulse chain: wf += ((1/hbar) F_g) a_g(w) (kernel_gp_compute.cuh:94-100,160-164) */
static void add_pulse_chain(const ws_t* w, int comp, size_t i, double wt, double* wr, double* wi) {
    const orc_params* p = w->p;
    for (int g = 0; g < p->src_groups[2]; g++) {
        double ar, ai;
        src_amp(p, 2, g, wt, &ar, &ai);
        const double* f = w->src[2][g][comp];
        const double tr = p->one_over_h_bar_s * f[2 * i], ti = p->one_over_h_bar_s * f[2 * i + 1];
        *wr += tr * ar - ti * ai;
        *wi += tr * ai + ti * ar;
    }
}

/* pump term of k_n: the single field at its amplitude, or the chain sum_g P_g Re a_g(w)
   (kernel_gp_compute.cuh:70-75) */
static double pump_term(const ws_t* w, int comp, size_t i, double wt, double pamp, double kn) {
    const orc_params* p = w->p;
    if (p->src_groups[0] > 0) {
        for (int g = 0; g < p->src_groups[0]; g++) {
            double ar, ai;
            src_amp(p, 0, g, wt, &ar, &ai);
            kn += w->src[0][g][comp][i] * ar;
        }
        return kn;
    }
    if (p->use_pump) kn += w->pump[comp][i] * pamp;
    return kn;
}

/* complex-reservoir pump term: k_n += sum_g P_g a_g(w) with the full complex amplitude
   (kernel_gp_compute.cuh:70-75; the single-field form has a real amplitude) */
static void pump_term_c(const ws_t* w, int comp, size_t i, double wt, double pamp, double* knr, double* kni) {
    const orc_params* p = w->p;
    if (p->src_groups[0] > 0) {
        for (int g = 0; g < p->src_groups[0]; g++) {
            double ar, ai;
            src_amp(p, 0, g, wt, &ar, &ai);
            const double v = w->src[0][g][comp][i];
            *knr += v * ar;
            *kni += v * ai;
        }
        return;
    }
    if (p->use_pump) *knr += w->pump[comp][i] * pamp;
}

/* complex-reservoir terms of one component: k_psi += (1/hbar) g_r n (-i psi) + 1/2 R n psi,
   k_n = -(gamma_r + R |psi|^2) n + pump (+ R n / dV with noise), n = (nr, ni) complex
   (kernel_gp_compute.cuh:66-82,166-181) */
static void cres_terms(const ws_t* w, int comp, size_t i, size_t ki, double wt, double pamp, int stoch, double nr, double ni,
                       double pr, double pi, double nrm, double* wr, double* wi) {
    const orc_params* p = w->p;
    const double ohb = p->one_over_h_bar_s;
    const double mr = MI_RE(pr, pi), mi = MI_IM(pr, pi);
    const double gr = ohb * p->g_r;
    *wr += (gr * nr) * mr - (gr * ni) * mi;
    *wi += (gr * nr) * mi + (gr * ni) * mr;
    const double hr = 0.5 * p->R;
    *wr += (hr * nr) * pr - (hr * ni) * pi;
    *wi += (hr * nr) * pi + (hr * ni) * pr;
    const double d = -(p->gamma_r + p->R * nrm);
    double knr = d * nr, kni = d * ni;
    pump_term_c(w, comp, i, wt, pamp, &knr, &kni);
    if (stoch) {
        knr += p->R * nr / p->dV;
        kni += p->R * ni / p->dV;
    }
    w->kn[comp][2 * ki] = knr;
    w->kn[comp][2 * ki + 1] = kni;
}

static void calc_k(ws_t* w, int index, double wt, int from_buffer) {
    const orc_params* p = w->p;
    const int ch = w->h - index;
    const int ro = w->ro;
    const double ohb = p->one_over_h_bar_s;
    const double pamp = p->pump_amp * (1.0 - wt) + wt * p->pump_amp_next;
    const double vamp = p->pot_amp * (1.0 - wt) + wt * p->pot_amp_next;
    const int stoch = p->use_stochastic && p->stochastic_amplitude > 0.0;
    #pragma omp parallel for schedule(static)
    for (int g = 0; g < w->S; g++) {
        size_t base = (size_t)g * w->N2h;
        size_t kbase = ((size_t)(index - 1) * w->S + g) * w->N2h;
        for (int r = w->h - ch; r < w->h + w->sgr + ch; r++)
            for (int c = w->h - ch; c < w->h + w->sgc + ch; c++) {
                size_t i = base + (size_t)r * ro + c;
                size_t ki = kbase + (size_t)r * ro + c;
                if (!p->tetm) {
                    const double* in = from_buffer ? w->bpsi[0] : w->psi[0];
                    const double* inn = from_buffer ? w->bn[0] : w->n[0];
                    const double pr = in[2 * i], pi = in[2 * i + 1];
                    const double mr = MI_RE(pr, pi), mi = MI_IM(pr, pi);
                    const double nrm = pr * pr + pi * pi;
                    /* m_eff_scaled * (C*psi + (N+S)/dy^2 + (E+W)/dx^2) */
                    double hr = p->m_eff_scaled * (p->m2_over_dx2_p_dy2 * pr + (in[2 * (i + ro)] + in[2 * (i - ro)]) * p->one_over_dy2 + (in[2 * (i + 1)] + in[2 * (i - 1)]) * p->one_over_dx2);
                    double hi = p->m_eff_scaled * (p->m2_over_dx2_p_dy2 * pi + (in[2 * (i + ro) + 1] + in[2 * (i - ro) + 1]) * p->one_over_dy2 + (in[2 * (i + 1) + 1] + in[2 * (i - 1) + 1]) * p->one_over_dx2);
                    double wr = MI_RE(hr, hi) * ohb, wi = MI_IM(hr, hi) * ohb;
                    wr += ohb * p->g_c * nrm * mr;
                    wi += ohb * p->g_c * nrm * mi;
                    wr -= 0.5 * p->gamma_c * pr;
                    wi -= 0.5 * p->gamma_c * pi;
                    if (p->use_reservoir && w->nel == 2) {
                        cres_terms(w, 0, i, ki, wt, pamp, stoch, inn[2 * i], inn[2 * i + 1], pr, pi, nrm, &wr, &wi);
                    } else if (p->use_reservoir) {
                        const double nv = inn[i];
                        wr += ohb * p->g_r * nv * mr;
                        wi += ohb * p->g_r * nv * mi;
                        wr += 0.5 * p->R * nv * pr;
                        wi += 0.5 * p->R * nv * pi;
                        double kn = -(p->gamma_r + p->R * nrm) * nv;
                        kn = pump_term(w, 0, i, wt, pamp, kn);
                        if (stoch) kn += p->R * nv / p->dV;
                        w->kn[0][ki] = kn;
                    }
                    if (p->src_groups[1] > 0) {
                        add_pot_chain(w, 0, i, wt, mr, mi, &wr, &wi);
                    } else if (p->use_potential) {
                        const double v = w->pot[0][i] * vamp;
                        wr += ohb * v * mr;
                        wi += ohb * v * mi;
                    }
                    add_pulse_chain(w, 0, i, wt, &wr, &wi);
                    if (stoch) {
                        wr -= ohb * p->g_c * mr / p->dV;
                        wi -= ohb * p->g_c * mi / p->dV;
                    }
                    w->kpsi[0][2 * ki] = wr;
                    w->kpsi[0][2 * ki + 1] = wi;
                } else {
                    /* TE-TM branch (kernel_gp_compute.cuh:112-235) */
                    const double* P = from_buffer ? w->bpsi[0] : w->psi[0];
                    const double* M = from_buffer ? w->bpsi[1] : w->psi[1];
                    const double* NP = from_buffer ? w->bn[0] : w->n[0];
                    const double* NM = from_buffer ? w->bn[1] : w->n[1];
                    const double idx2 = p->one_over_dx2, idy2 = p->one_over_dy2;
                    double horp_r = (P[2 * (i + 1)] + P[2 * (i - 1)]) * idx2, horp_i = (P[2 * (i + 1) + 1] + P[2 * (i - 1) + 1]) * idx2;
                    double verp_r = (P[2 * (i + ro)] + P[2 * (i - ro)]) * idy2, verp_i = (P[2 * (i + ro) + 1] + P[2 * (i - ro) + 1]) * idy2;
                    double horm_r = (M[2 * (i + 1)] + M[2 * (i - 1)]) * idx2, horm_i = (M[2 * (i + 1) + 1] + M[2 * (i - 1) + 1]) * idx2;
                    double verm_r = (M[2 * (i + ro)] + M[2 * (i - ro)]) * idy2, verm_i = (M[2 * (i + ro) + 1] + M[2 * (i - ro) + 1]) * idy2;
                    /* diagonal combination d = f[i+ro+1] + f[i-ro-1] - f[i+ro-1] - f[i-ro+1] */
                    double dm_r = M[2 * (i + ro + 1)] + M[2 * (i - ro - 1)] - M[2 * (i + ro - 1)] - M[2 * (i - ro + 1)];
                    double dm_i = M[2 * (i + ro + 1) + 1] + M[2 * (i - ro - 1) + 1] - M[2 * (i + ro - 1) + 1] - M[2 * (i - ro + 1) + 1];
                    double dp_r = P[2 * (i + ro + 1)] + P[2 * (i - ro - 1)] - P[2 * (i + ro - 1)] - P[2 * (i - ro + 1)];
                    double dp_i = P[2 * (i + ro + 1) + 1] + P[2 * (i - ro - 1) + 1] - P[2 * (i + ro - 1) + 1] - P[2 * (i - ro + 1) + 1];
                    const double hdd = 0.5 / p->dx / p->dy; /* (i/2)/dx/dy */
                    /* (i*hdd) * d = (-hdd*d_i, hdd*d_r) */
                    double crp_r = horm_r - verm_r + (-hdd * dm_i), crp_i = horm_i - verm_i + hdd * dm_r;
                    double crm_r = horp_r - verp_r - (-hdd * dp_i), crm_i = horp_i - verp_i - hdd * dp_r;
                    const double ppr = P[2 * i], ppi = P[2 * i + 1], pmr = M[2 * i], pmi = M[2 * i + 1];
                    double regp_r = p->m2_over_dx2_p_dy2 * ppr + horp_r + verp_r, regp_i = p->m2_over_dx2_p_dy2 * ppi + horp_i + verp_i;
                    double regm_r = p->m2_over_dx2_p_dy2 * pmr + horm_r + verm_r, regm_i = p->m2_over_dx2_p_dy2 * pmi + horm_i + verm_i;
                    const int cres = w->nel == 2;
                    const double np_ = cres ? 0.0 : NP[i], nm_ = cres ? 0.0 : NM[i];
                    const double nrmp = ppr * ppr + ppi * ppi, nrmm = pmr * pmr + pmi * pmi;
                    for (int comp = 0; comp < 2; comp++) {
                        const double sr = comp ? pmr : ppr, si = comp ? pmi : ppi;
                        const double mr = MI_RE(sr, si), mi = MI_IM(sr, si);
                        const double rr = comp ? regm_r : regp_r, ri = comp ? regm_i : regp_i;
                        const double cr = comp ? crm_r : crp_r, ci = comp ? crm_i : crp_i;
                        const double nv = comp ? nm_ : np_;
                        const double nself = comp ? nrmm : nrmp, nother = comp ? nrmp : nrmm;
                        double res_r = ohb * p->m_eff_scaled * MI_RE(rr, ri);
                        double res_i = ohb * p->m_eff_scaled * MI_IM(rr, ri);
                        if (p->src_groups[1] > 0) {
                            add_pot_chain(w, comp, i, wt, mr, mi, &res_r, &res_i);
                        } else if (p->use_potential) {
                            const double v = w->pot[comp][i] * vamp;
                            res_r += ohb * v * mr;
                            res_i += ohb * v * mi;
                        }
                        res_r += ohb * p->g_c * nself * mr;  res_i += ohb * p->g_c * nself * mi;
                        if (!cres) {
                            res_r += ohb * p->g_r * nv * mr;     res_i += ohb * p->g_r * nv * mi;
                            res_r += 0.5 * p->R * nv * sr;       res_i += 0.5 * p->R * nv * si;
                        } else {
                            const double* NC = comp ? NM : NP;
                            cres_terms(w, comp, i, ki, wt, pamp, stoch, NC[2 * i], NC[2 * i + 1], sr, si, nself, &res_r, &res_i);
                        }
                        res_r -= 0.5 * p->gamma_c * sr;      res_i -= 0.5 * p->gamma_c * si;
                        res_r += ohb * p->g_pm * nother * mr; res_i += ohb * p->g_pm * nother * mi;
                        res_r += ohb * p->delta_LT * MI_RE(cr, ci);
                        res_i += ohb * p->delta_LT * MI_IM(cr, ci);
                        add_pulse_chain(w, comp, i, wt, &res_r, &res_i);
                        if (stoch) {
                            res_r -= ohb * p->g_c * mr / p->dV;
                            res_i -= ohb * p->g_c * mi / p->dV;
                        }
                        w->kpsi[comp][2 * ki] = res_r;
                        w->kpsi[comp][2 * ki + 1] = res_i;
                        if (!cres) {
                            double kn = -(p->gamma_r + p->R * nself) * nv;
                            kn = pump_term(w, comp, i, wt, pamp, kn);
                            if (stoch) kn += p->R * nv / p->dV;
                            w->kn[comp][ki] = kn;
                        }
                    }
                }
            }
    }
}

/* sum_{j} weight_j k_slot_j with zero weights skipped (kernel_summation.cuh:7-29),
   left-to-right accumulation. */
static inline void ksum(const ws_t* w, const double* kf, int elem, size_t off, int nw, const int* slot, const double* wts, double* out) {
    out[0] = 0.0; if (elem == 2) out[1] = 0.0;
    int first = 1;
    for (int j = 0; j < nw; j++) {
        if (wts[j] == 0.0) continue;
        size_t o = ((size_t)slot[j] * w->S * w->N2h + off) * elem;
        if (first) {
            out[0] = wts[j] * kf[o];
            if (elem == 2) out[1] = wts[j] * kf[o + 1];
            first = 0;
        } else {
            out[0] += wts[j] * kf[o];
            if (elem == 2) out[1] += wts[j] * kf[o + 1];
        }
    }
}

/* Stochastic increment (kernel_summation.cuh:209-225): w_sum * xi * sqrt((R n + gamma_c)/(4 dV)),
   n = step-start reservoir+ (used for both components). */
static void add_dw(const ws_t* w, size_t i, double wsum, double* out_r, double* out_i) {
    const orc_params* p = w->p;
    if (!(p->use_stochastic && p->stochastic_amplitude > 0.0)) return;
    if (p->use_reservoir && w->nel == 2) {
        /* complex reservoir: principal square root of (R n + gamma_c) / (4 dV), times w_sum xi */
        const double complex z = ((p->R * w->n[0][2 * i] + p->gamma_c) + I * (p->R * w->n[0][2 * i + 1])) / (4.0 * p->dV);
        const double complex sq = csqrt(z);
        const double xr = wsum * w->xi[2 * i], xi = wsum * w->xi[2 * i + 1];
        *out_r += xr * creal(sq) - xi * cimag(sq);
        *out_i += xr * cimag(sq) + xi * creal(sq);
        return;
    }
    double s = p->use_reservoir ? sqrt((p->R * w->n[0][i] + p->gamma_c) / (4.0 * p->dV)) : sqrt(p->gamma_c / (4.0 * p->dV));
    *out_r += wsum * w->xi[2 * i] * s;
    *out_i += wsum * w->xi[2 * i + 1] * s;
}

/* INTERMEDIATE_SUM_K(index, weights) (cuda_macro.cuh:153-182): buffer = X + dt*sum
   on extent halo-index; single weight -> slot index-1, else slots 0..n-1. */
static void sum_to_buffer(ws_t* w, int index, int nw, const double* wts) {
    const int ch = w->h - index;
    const int imag = w->p->imag_time_amplitude != 0.0;
    int slot[13];
    for (int j = 0; j < nw; j++) slot[j] = (nw == 1) ? index - 1 : j;
    double wsum = 0.0;
    for (int j = 0; j < nw; j++) wsum += wts[j];
    const double dt = w->dt;
    #pragma omp parallel for schedule(static)
    for (int g = 0; g < w->S; g++)
        for (int r = w->h - ch; r < w->h + w->sgr + ch; r++)
            for (int c = w->h - ch; c < w->h + w->sgc + ch; c++) {
                size_t i = (size_t)g * w->N2h + (size_t)r * w->ro + c;
                for (int comp = 0; comp < w->ncomp; comp++) {
                    double s[2];
                    ksum(w, w->kpsi[comp], 2, i, nw, slot, wts, s);
                    /* complex_dt: input + (0, -dt) * sum (kernel_summation.cuh:195-201) */
                    double orr = imag ? w->psi[comp][2 * i] + dt * s[1] : w->psi[comp][2 * i] + dt * s[0];
                    double oi = imag ? w->psi[comp][2 * i + 1] - dt * s[0] : w->psi[comp][2 * i + 1] + dt * s[1];
                    add_dw(w, i, wsum, &orr, &oi);
                    w->bpsi[comp][2 * i] = orr;
                    w->bpsi[comp][2 * i + 1] = oi;
                    if (w->p->use_reservoir && w->nel == 2) {
                        /* complex reservoir: input + dt sum, or + (0, -dt) sum in imaginary time */
                        ksum(w, w->kn[comp], 2, i, nw, slot, wts, s);
                        const double* x = w->n[comp];
                        w->bn[comp][2 * i] = imag ? x[2 * i] + dt * s[1] : x[2 * i] + dt * s[0];
                        w->bn[comp][2 * i + 1] = imag ? x[2 * i + 1] - dt * s[0] : x[2 * i + 1] + dt * s[1];
                    } else if (w->p->use_reservoir) {
                        ksum(w, w->kn[comp], 1, i, nw, slot, wts, s);
                        w->bn[comp][i] = w->n[comp][i] + dt * s[0];
                    }
                }
            }
}

/* FINAL_SUM_K(index, weights) (cuda_macro.cuh:200-229): X += dt*sum b_j k_j on the whole
   padded region; single weight -> slot index-1, else slots 0..n-1. */
static void final_sum(ws_t* w, int index, int nw, const double* wts) {
    const int ch = w->h;
    const int imag = w->p->imag_time_amplitude != 0.0;
    int slot[13];
    for (int j = 0; j < nw; j++) slot[j] = (nw == 1) ? index - 1 : j;
    double wsum = 0.0;
    for (int j = 0; j < nw; j++) wsum += wts[j];
    const double dt = w->dt;
    #pragma omp parallel for schedule(static)
    for (int g = 0; g < w->S; g++)
        for (int r = w->h - ch; r < w->h + w->sgr + ch; r++)
            for (int c = w->h - ch; c < w->h + w->sgc + ch; c++) {
                size_t i = (size_t)g * w->N2h + (size_t)r * w->ro + c;
                /* the noise amplitude reads the step-start reservoir: evaluate before updating n */
                double dwr = 0.0, dwi = 0.0;
                add_dw(w, i, wsum, &dwr, &dwi);
                for (int comp = 0; comp < w->ncomp; comp++) {
                    double s[2];
                    ksum(w, w->kpsi[comp], 2, i, nw, slot, wts, s);
                    if (imag) { /* (0, -dt) * sum (kernel_summation.cuh:227-233) */
                        w->psi[comp][2 * i] += dt * s[1];
                        w->psi[comp][2 * i + 1] -= dt * s[0];
                    } else {
                        w->psi[comp][2 * i] += dt * s[0];
                        w->psi[comp][2 * i + 1] += dt * s[1];
                    }
                    w->psi[comp][2 * i] += dwr;
                    w->psi[comp][2 * i + 1] += dwi;
                }
                if (w->p->use_reservoir)
                    for (int comp = 0; comp < w->ncomp; comp++) {
                        double s[2];
                        if (w->nel == 2) {
                            ksum(w, w->kn[comp], 2, i, nw, slot, wts, s);
                            if (imag) {
                                w->n[comp][2 * i] += dt * s[1];
                                w->n[comp][2 * i + 1] -= dt * s[0];
                            } else {
                                w->n[comp][2 * i] += dt * s[0];
                                w->n[comp][2 * i + 1] += dt * s[1];
                            }
                            continue;
                        }
                        ksum(w, w->kn[comp], 1, i, nw, slot, wts, s);
                        w->n[comp][i] += dt * s[0];
                    }
            }
}

static void sync_state(ws_t* w) {
    for (int c = 0; c < w->ncomp; c++) {
        sync_halos(w, w->psi[c], 2);
        if (w->p->use_reservoir) sync_halos(w, w->n[c], w->nel);
    }
}

/* Own noise stream for one step: xi = amp*sqrt(dt)*N2(seed, step, global cell), 0 outside a zero-BC domain. */
static void gen_noise(ws_t* w, int64_t step, double dt) {
    const orc_params* p = w->p;
    if (!(p->use_stochastic && p->stochastic_amplitude > 0.0)) return;
    const double a = p->stochastic_amplitude * sqrt(dt);
    #pragma omp parallel for schedule(static)
    for (int g = 0; g < w->S; g++)
        for (int tr = 0; tr < w->sgr + 2 * w->h; tr++)
            for (int tc = 0; tc < w->sgc + 2 * w->h; tc++) {
                size_t i = (size_t)g * w->N2h + (size_t)tr * w->ro + tc;
                long gc = global_cell(w, g, tr, tc);
                double re = 0.0, im = 0.0;
                if (gc >= 0) orc_noise_normal2(p->noise_seed, (uint64_t)step, (uint32_t)gc, 0, &re, &im);
                w->xi[2 * i] = a * re;
                w->xi[2 * i + 1] = a * im;
            }
}

static void load_state(ws_t* w, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                       const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus) {
    const orc_params* p = w->p;
    to_sub(w, w->psi[0], psi_plus, 2);
    if (p->use_reservoir) to_sub(w, w->n[0], n_plus, w->nel);
    /* (a source given as a chain has no single field: NULL) */
    if (p->use_pump && pump_plus) { to_sub(w, w->pump[0], pump_plus, 1); sync_halos(w, w->pump[0], 1); }
    if (p->use_potential && pot_plus) { to_sub(w, w->pot[0], pot_plus, 1); sync_halos(w, w->pot[0], 1); }
    for (int q = 0; q < 3; q++)
        for (int g = 0; g < p->src_groups[q]; g++)
            for (int c = 0; c < w->ncomp; c++) {
                const int el = q == 2 ? 2 : 1;
                if (!p->src_field[q][g][c]) continue;
                to_sub(w, w->src[q][g][c], p->src_field[q][g][c], el);
                sync_halos(w, w->src[q][g][c], el);
            }
    if (p->tetm) {
        to_sub(w, w->psi[1], psi_minus, 2);
        if (p->use_reservoir) to_sub(w, w->n[1], n_minus, w->nel);
        if (p->use_pump && pump_minus) { to_sub(w, w->pump[1], pump_minus, 1); sync_halos(w, w->pump[1], 1); }
        if (p->use_potential && pot_minus) { to_sub(w, w->pot[1], pot_minus, 1); sync_halos(w, w->pot[1], 1); }
    }
}

static void store_state(ws_t* w, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus, int from_buffer) {
    const orc_params* p = w->p;
    to_full(w, from_buffer ? w->bpsi[0] : w->psi[0], psi_plus, 2);
    if (p->use_reservoir) to_full(w, from_buffer ? w->bn[0] : w->n[0], n_plus, w->nel);
    if (p->tetm) {
        to_full(w, from_buffer ? w->bpsi[1] : w->psi[1], psi_minus, 2);
        if (p->use_reservoir) to_full(w, from_buffer ? w->bn[1] : w->n[1], n_minus, w->nel);
    }
}

/* ------------------------------------------------------------------------ */
/* RK4 (runge_kutta.cu:74-94)                                               */
/* ------------------------------------------------------------------------ */
static void rk4_step(ws_t* w) {
    static const double half = 0.5, one = 1.0;
    static const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    sync_state(w);
    calc_k(w, 1, 0.0, 0);
    sum_to_buffer(w, 1, 1, &half);
    calc_k(w, 2, 0.5, 1);
    sum_to_buffer(w, 2, 1, &half);
    calc_k(w, 3, 0.5, 1);
    sum_to_buffer(w, 3, 1, &one);
    calc_k(w, 4, 1.0, 1);
    final_sum(w, 4, 4, b);
}

int orc_rk4_run(const orc_params* p, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                double dt, int steps, int64_t step0) {
    ws_t w;
    if (ws_init(&w, p, 4)) return -1;
    w.dt = dt;
    load_state(&w, psi_plus, psi_minus, n_plus, n_minus, pump_plus, pump_minus, pot_plus, pot_minus);
    for (int s = 0; s < steps; s++) {
        gen_noise(&w, step0 + s, dt);
        rk4_step(&w);
    }
    store_state(&w, psi_plus, psi_minus, n_plus, n_minus, 0);
    ws_free(&w);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Dormand-Prince 5(4) (dormand_prince.cu:37-95)                            */
/* ------------------------------------------------------------------------ */
static const double DP_A2[1] = {1.0 / 5.0};
static const double DP_A3[2] = {3.0 / 40.0, 9.0 / 40.0};
static const double DP_A4[3] = {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0};
static const double DP_A5[4] = {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0};
static const double DP_A6[5] = {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0};
static const double DP_B[6] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0};
static const double DP_E[7] = {35.0 / 384.0 - 5179.0 / 57600.0, 0.0, 500.0 / 1113.0 - 7571.0 / 16695.0, 125.0 / 192.0 - 393.0 / 640.0,
                               -2187.0 / 6784.0 + 92097.0 / 339200.0, 11.0 / 84.0 - 187.0 / 2100.0, -1.0 / 40.0};

static void dp45_stages(ws_t* w) {
    sync_state(w);
    calc_k(w, 1, 0.0, 0);
    sum_to_buffer(w, 1, 1, DP_A2);
    calc_k(w, 2, 1.0 / 5.0, 1);
    sum_to_buffer(w, 2, 2, DP_A3);
    calc_k(w, 3, 3.0 / 10.0, 1);
    sum_to_buffer(w, 3, 3, DP_A4);
    calc_k(w, 4, 4.0 / 5.0, 1);
    sum_to_buffer(w, 4, 4, DP_A5);
    calc_k(w, 5, 8.0 / 9.0, 1);
    sum_to_buffer(w, 5, 5, DP_A6);
    calc_k(w, 6, 1.0, 1);
    sum_to_buffer(w, 6, 6, DP_B);
    calc_k(w, 7, 1.0, 1);
}

int orc_dp45_attempt(const orc_params* p, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                     const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                     double dt, double noise_dt, int64_t step_index, double* err_sum, double* norm) {
    ws_t w;
    if (ws_init(&w, p, 7)) return -1;
    w.dt = dt;
    load_state(&w, psi_plus, psi_minus, n_plus, n_minus, pump_plus, pump_minus, pot_plus, pot_minus);
    gen_noise(&w, step_index, noise_dt);
    dp45_stages(&w);
    /* ERROR_K (runge_sum_to_error, kernel_summation.cuh:270-289) on interior cells;
       norm = sum |buffer psi+|^2 on interior cells. */
    double es = 0.0, ns = 0.0;
    int slot[7] = {0, 1, 2, 3, 4, 5, 6};
    for (int g = 0; g < w.S; g++)
        for (int r = w.h; r < w.h + w.sgr; r++)
            for (int c = w.h; c < w.h + w.sgc; c++) {
                size_t i = (size_t)g * w.N2h + (size_t)r * w.ro + c;
                double s[2];
                ksum(&w, w.kpsi[0], 2, i, 7, slot, DP_E, s);
                double er = dt * s[0], ei = dt * s[1];
                double e = er * er + ei * ei;
                es += p->tetm ? 2.0 * e : e;
                ns += w.bpsi[0][2 * i] * w.bpsi[0][2 * i] + w.bpsi[0][2 * i + 1] * w.bpsi[0][2 * i + 1];
            }
    *err_sum = es;
    *norm = ns;
    store_state(&w, psi_plus, psi_minus, n_plus, n_minus, 1);
    ws_free(&w);
    return 0;
}

/* Solver::adaptTimeStep (solver.cu:72-105), power 1/4. */
int orc_adapt_dt(double err_sum, double norm, double tolerance, double dt, double dt_min, double dt_max, double* new_dt, double* final_error) {
    double fe = fabs(err_sum / norm);
    double dh_arg = tolerance / 2.0 / (fe > DBL_MIN ? fe : DBL_MIN);
    double dh = pow(dh_arg, 0.25);
    if (isnan(dh)) { dh = 0.9; fe = DBL_MAX; }
    if (isnan(fe)) { dh = 0.9; fe = DBL_MAX; }
    double nd = dt * dh < dt_max ? dt * dh : dt_max;
    if (nd < dt_min) nd = dt_min;
    *new_dt = nd;
    *final_error = fe;
    return fe < tolerance;
}

int64_t orc_dp45_run(const orc_params* p, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                     const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                     double* t, double* dt, double t_max, double tolerance, double dt_min, double dt_max,
                     int max_rejections, int64_t* n_rejected) {
    const size_t N = (size_t)p->N_c * p->N_r;
    const int nc = p->tetm ? 2 : 1;
    double *ps[2] = {psi_plus, psi_minus}, *ns[2] = {n_plus, n_minus};
    double *tp[2] = {0, 0}, *tn[2] = {0, 0};
    for (int c = 0; c < nc; c++) {
        tp[c] = malloc(2 * N * sizeof(double));
        tn[c] = malloc((p->complex_reservoir ? 2 : 1) * N * sizeof(double));
    }
    int64_t accepted = 0, rejected = 0;
    int64_t step = 0;
    while (*t < t_max) {
        const double noise_dt = *dt; /* noise drawn once per iterate() with sqrt(dt) (solver.cu:36-45) */
        *t = *t + *dt; /* solver.cu:50, before the attempts */
        int rej = 0;
        for (;;) {
            for (int c = 0; c < nc; c++) {
                memcpy(tp[c], ps[c], 2 * N * sizeof(double));
                if (p->use_reservoir) memcpy(tn[c], ns[c], (p->complex_reservoir ? 2 : 1) * N * sizeof(double));
            }
            double es, nm, nd, fe;
            if (orc_dp45_attempt(p, tp[0], tp[1], tn[0], tn[1], pump_plus, pump_minus, pot_plus, pot_minus, *dt, noise_dt, step, &es, &nm)) return -1;
            int acc = orc_adapt_dt(es, nm, tolerance, *dt, dt_min, dt_max, &nd, &fe);
            *dt = nd;
            if (acc) break;
            rejected++;
            if (++rej > max_rejections) {
                *n_rejected = rejected;
                for (int c = 0; c < nc; c++) { free(tp[c]); free(tn[c]); }
                return -2;
            }
        }
        for (int c = 0; c < nc; c++) {
            memcpy(ps[c], tp[c], 2 * N * sizeof(double));
            if (p->use_reservoir) memcpy(ns[c], tn[c], (p->complex_reservoir ? 2 : 1) * N * sizeof(double));
        }
        accepted++;
        step++;
    }
    for (int c = 0; c < nc; c++) { free(tp[c]); free(tn[c]); }
    *n_rejected = rejected;
    return accepted;
}

/* ------------------------------------------------------------------------ */
/* Generic iterators: the step() sequences of solver/iterator/<name>.cu as data   */
/* ------------------------------------------------------------------------ */
/* One macro call of a step(): 'K' CALCULATE_K(index, w, input = buffer?),
   'I' INTERMEDIATE_SUM_K(index, weights), 'F' FINAL_SUM_K(index, weights),
   'E' ERROR_K(weights) (the psi+ error and the buffer norm on interior cells). */
typedef struct {
    char op;
    int index;
    double w;
    int nw;
    double wts[13];
} orc_op;

typedef struct {
    const char* key;
    int halo;        /* halo_size_ of the reference iterator */
    double power;    /* adaptTimeStep exponent; 0: no variable time step */
    int nfixed, nadapt;
    orc_op fixed[28];
    orc_op adapt[28];
} orc_iter;

#define K1_ {'K', 1, 0.0, 0, {0}}
#define KB(i, w) {'K', i, w, 0, {0}}

static const orc_iter ITERS[] = {
    /* Newton::step (newton.cu:24-31) */
    {"Newton", 1, 0, 2, 0, {K1_, {'F', 1, 0, 1, {1.0}}}, {{0}}},
    /* MidPoint::step (explicit_mp.cu:28-39); halo_size_ = 3 (explicit_mp.cu:15) */
    {"MP", 3, 0, 4, 0, {K1_, {'I', 1, 0, 1, {0.5}}, KB(2, 0.5), {'F', 2, 0, 2, {0.0, 1.0}}}, {{0}}},
    /* Heun2::step (heun.cu:31-42) */
    {"Heun2", 2, 0, 4, 0, {K1_, {'I', 1, 0, 1, {1.0}}, KB(2, 1.0), {'F', 3, 0, 2, {0.5, 0.5}}}, {{0}}},
    /* Ralston::step (ralston.cu:31-42) */
    {"Ralston", 2, 0, 4, 0, {K1_, {'I', 1, 0, 1, {2.0 / 3.0}}, KB(2, 2.0 / 3.0), {'F', 2, 0, 2, {1.0 / 4.0, 3.0 / 4.0}}}, {{0}}},
    /* Heun3::step (heun.cu:66-82) */
    {"Heun3", 3, 0, 6, 0,
     {K1_, {'I', 1, 0, 1, {1.0 / 3.0}}, KB(2, 1.0 / 3.0), {'I', 2, 0, 2, {0.0, 2.0 / 3.0}}, KB(3, 2.0 / 3.0),
      {'F', 3, 0, 3, {1.0 / 4.0, 0.0, 3.0 / 4.0}}},
     {{0}}},
    /* Ralston3::step (ralston.cu:66-81) */
    {"Ralston3", 3, 0, 6, 0,
     {K1_, {'I', 1, 0, 1, {1.0 / 2.0}}, KB(2, 1.0 / 2.0), {'I', 2, 0, 2, {0.0, 3.0 / 4.0}}, KB(3, 3.0 / 4.0),
      {'F', 3, 0, 3, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}}},
     {{0}}},
    /* RungeKutta3::step (runge_kutta.cu:33-48); its final sum names a fourth weight for a K
       slot the solver never writes (k_max_ = 3): the three-weight sum of its tableau string
       (runge_kutta.cu:23-30) is used */
    {"RK3", 3, 0, 6, 0,
     {K1_, {'I', 1, 0, 1, {0.5}}, KB(2, 0.5), {'I', 2, 0, 2, {-1.0, 2.0}}, KB(3, 1.0),
      {'F', 3, 0, 3, {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}}},
     {{0}}},
    /* SSPRK3::step (SSPRK3.cu:27-42), third stage at w = 1 as passed */
    {"SSPRK3", 3, 0, 6, 0,
     {K1_, {'I', 1, 0, 1, {1.0}}, KB(2, 1.0), {'I', 2, 0, 2, {1.0 / 4.0, 1.0 / 4.0}}, KB(3, 1.0),
      {'F', 3, 0, 3, {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}}},
     {{0}}},
    /* HouwenWray::step (houwen_wray.cu:27-42) */
    {"HW", 3, 0, 6, 0,
     {K1_, {'I', 1, 0, 1, {8.0 / 15.0}}, KB(2, 8.0 / 15.0), {'I', 2, 0, 2, {1.0 / 4.0, 5.0 / 12.0}}, KB(3, 2.0 / 3.0),
      {'F', 3, 0, 3, {1.0 / 4.0, 0.0, 3.0 / 4.0}}},
     {{0}}},
    /* RungeKutta4::step (runge_kutta.cu:74-94) */
    {"RK4", 4, 0, 8, 0,
     {K1_, {'I', 1, 0, 1, {0.5}}, KB(2, 0.5), {'I', 2, 0, 1, {0.5}}, KB(3, 0.5), {'I', 3, 0, 1, {1.0}}, KB(4, 1.0),
      {'F', 4, 0, 4, {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0}}},
     {{0}}},
    /* Ralston4::step (ralston.cu:106-125); the reference runs its four stages with
       k_max_ = halo_size_ = 3 (ralston.cu:92-93), which leaves the fourth K outside the K
       buffer: here with the halo of four that the four stages need */
    {"Ralston4", 4, 0, 8, 0,
     {K1_, {'I', 1, 0, 1, {0.4}}, KB(2, 0.4), {'I', 2, 0, 2, {0.29697760924, 0.15875964497}}, KB(3, 0.51823725421),
      {'I', 3, 0, 3, {0.21810038822, -3.05096514869, 3.83286476047}}, KB(4, 0.51823725421),
      {'F', 4, 0, 4, {0.17476028226, -0.55148066287, 1.2055355994, 0.17118478122}}},
     {{0}}},
    /* Rule38::step (rule38.cu:28-48) */
    {"Rule38", 4, 0, 8, 0,
     {K1_, {'I', 1, 0, 1, {1.0 / 3.0}}, KB(2, 1.0 / 3.0), {'I', 2, 0, 2, {-1.0 / 3.0, 1.0}}, KB(3, 2.0 / 3.0),
      {'I', 3, 0, 3, {1.0, -1.0, 1.0}}, KB(4, 1.0), {'F', 4, 0, 4, {1.0 / 8.0, 3.0 / 8.0, 3.0 / 8.0, 1.0 / 8.0}}},
     {{0}}},
    /* Nystroem::step (nystroem.cu:30-57) */
    {"Nystroem", 6, 0, 12, 0,
     {K1_, {'I', 1, 0, 1, {1.0 / 3.0}}, KB(2, 1.0 / 3.0), {'I', 2, 0, 2, {4.0 / 25.0, 6.0 / 25.0}}, KB(3, 2.0 / 5.0),
      {'I', 3, 0, 3, {1.0 / 4.0, -3.0, 15.0 / 4.0}}, KB(4, 1.0),
      {'I', 4, 0, 4, {2.0 / 27.0, 10.0 / 9.0, -50.0 / 81.0, 8.0 / 81.0}}, KB(5, 2.0 / 3.0),
      {'I', 5, 0, 4, {2.0 / 25.0, 12.0 / 25.0, 2.0 / 15.0, 8.0 / 75.0}}, KB(6, 4.0 / 5.0),
      {'F', 6, 0, 6, {23.0 / 192.0, 0.0, 125.0 / 192.0, 0.0, -27.0 / 64.0, 125.0 / 192.0}}},
     {{0}}},
    /* Fehlberg2::step (fehlberg.cu:33-70) */
    {"Fehlberg2", 3, 0.25, 6, 7,
     {K1_, {'I', 1, 0, 1, {1.0 / 2.0}}, KB(2, 1.0 / 2.0), {'I', 2, 0, 2, {1.0 / 256.0, 255.0 / 256.0}}, KB(3, 1.0),
      {'F', 3, 0, 3, {1.0 / 512.0, 255.0 / 256.0, 1.0 / 512.0}}},
     {K1_, {'I', 1, 0, 1, {1.0 / 2.0}}, KB(2, 1.0 / 2.0), {'I', 2, 0, 2, {1.0 / 256.0, 255.0 / 256.0}}, KB(3, 1.0),
      {'I', 3, 0, 3, {1.0 / 512.0, 255.0 / 256.0, 1.0 / 512.0}}, {'E', 3, 0, 3, {1.0 / 512.0 - 1.0 / 256.0, 0.0, 1.0 / 512.0}}}},
    /* BogackiShampine::step (bogacki.cu:28-75); a32 = 3/9 as written there (bogacki.cu:41) */
    {"BoSha", 4, 1.0 / 3.0, 6, 9,
     {K1_, {'I', 1, 0, 1, {1.0 / 2.0}}, KB(2, 1.0 / 2.0), {'I', 2, 0, 2, {0.0, 3.0 / 9.0}}, KB(3, 3.0 / 4.0),
      {'F', 3, 0, 3, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}}},
     {K1_, {'I', 1, 0, 1, {1.0 / 2.0}}, KB(2, 1.0 / 2.0), {'I', 2, 0, 2, {0.0, 3.0 / 9.0}}, KB(3, 3.0 / 4.0),
      {'I', 3, 0, 3, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}}, KB(4, 1.0),
      {'E', 4, 0, 4, {2.0 / 9.0 - 7.0 / 24.0, 1.0 / 3.0 - 1.0 / 4.0, 4.0 / 9.0 - 1.0 / 3.0, 0.0 - 1.0 / 8.0}},
      {'I', 3, 0, 3, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}}}},
    /* CashKarp::step (cash_karp.cu:31-83); a41 = 3/1 (cash_karp.cu:47) and the K5 error
       weight -2187/6784 (cash_karp.cu:70) as written there */
    {"CashKarp", 6, 0.25, 12, 13,
     {K1_, {'I', 1, 0, 1, {1.0 / 5.0}}, KB(2, 1.0 / 5.0), {'I', 2, 0, 2, {3.0 / 40.0, 9.0 / 40.0}}, KB(3, 3.0 / 10.0),
      {'I', 3, 0, 3, {3.0 / 1.0, -9.0 / 10.0, 6.0 / 5.0}}, KB(4, 3.0 / 5.0),
      {'I', 4, 0, 4, {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0}}, KB(5, 1.0),
      {'I', 5, 0, 5, {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}}, KB(6, 7.0 / 8.0),
      {'F', 6, 0, 6, {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0}}},
     {K1_, {'I', 1, 0, 1, {1.0 / 5.0}}, KB(2, 1.0 / 5.0), {'I', 2, 0, 2, {3.0 / 40.0, 9.0 / 40.0}}, KB(3, 3.0 / 10.0),
      {'I', 3, 0, 3, {3.0 / 1.0, -9.0 / 10.0, 6.0 / 5.0}}, KB(4, 3.0 / 5.0),
      {'I', 4, 0, 4, {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0}}, KB(5, 1.0),
      {'I', 5, 0, 5, {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}}, KB(6, 7.0 / 8.0),
      {'I', 6, 0, 6, {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0}},
      {'E', 6, 0, 6,
       {37.0 / 378.0 - 2825.0 / 27648.0, 0.0, 250.0 / 621.0 - 18575.0 / 48384.0, 125.0 / 594.0 - 13525.0 / 55296.0,
        0.0 - 2187.0 / 6784.0, 512.0 / 1771.0 - 1.0 / 4.0}}}},
    /* Fehlberg5::step (fehlberg.cu:100-152); the K5 error weight -9/50 - 1/5 as written
       there (fehlberg.cu:137) */
    {"Fehlberg5", 6, 0.25, 12, 13,
     {K1_, {'I', 1, 0, 1, {1.0 / 4.0}}, KB(2, 1.0 / 4.0), {'I', 2, 0, 2, {3.0 / 32.0, 9.0 / 32.0}}, KB(3, 3.0 / 8.0),
      {'I', 3, 0, 3, {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0}}, KB(4, 12.0 / 13.0),
      {'I', 4, 0, 4, {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0}}, KB(5, 1.0),
      {'I', 5, 0, 5, {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0}}, KB(6, 1.0 / 2.0),
      {'F', 6, 0, 6, {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0}}},
     {K1_, {'I', 1, 0, 1, {1.0 / 4.0}}, KB(2, 1.0 / 4.0), {'I', 2, 0, 2, {3.0 / 32.0, 9.0 / 32.0}}, KB(3, 3.0 / 8.0),
      {'I', 3, 0, 3, {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0}}, KB(4, 12.0 / 13.0),
      {'I', 4, 0, 4, {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0}}, KB(5, 1.0),
      {'I', 5, 0, 5, {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0}}, KB(6, 1.0 / 2.0),
      {'I', 6, 0, 6, {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0}},
      {'E', 6, 0, 6,
       {16.0 / 135.0 - 25.0 / 216.0, 0.0, 6656.0 / 12825.0 - 1408.0 / 2565.0, 28561.0 / 56430.0 - 2197.0 / 4104.0,
        -9.0 / 50.0 - 1.0 / 5.0, 2.0 / 55.0}}}},
    /* DormandPrince45::step (dormand_prince.cu:37-95) */
    {"DP45", 7, 0.25, 12, 14,
     {K1_, {'I', 1, 0, 1, {1.0 / 5.0}}, KB(2, 1.0 / 5.0), {'I', 2, 0, 2, {3.0 / 40.0, 9.0 / 40.0}}, KB(3, 3.0 / 10.0),
      {'I', 3, 0, 3, {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0}}, KB(4, 4.0 / 5.0),
      {'I', 4, 0, 4, {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0}}, KB(5, 8.0 / 9.0),
      {'I', 5, 0, 5, {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0}}, KB(6, 1.0),
      {'F', 6, 0, 6, {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}}},
     {K1_, {'I', 1, 0, 1, {1.0 / 5.0}}, KB(2, 1.0 / 5.0), {'I', 2, 0, 2, {3.0 / 40.0, 9.0 / 40.0}}, KB(3, 3.0 / 10.0),
      {'I', 3, 0, 3, {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0}}, KB(4, 4.0 / 5.0),
      {'I', 4, 0, 4, {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0}}, KB(5, 8.0 / 9.0),
      {'I', 5, 0, 5, {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0}}, KB(6, 1.0),
      {'I', 6, 0, 6, {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}}, KB(7, 1.0),
      {'E', 7, 0, 7,
       {35.0 / 384.0 - 5179.0 / 57600.0, 0.0, 500.0 / 1113.0 - 7571.0 / 16695.0, 125.0 / 192.0 - 393.0 / 640.0,
        -2187.0 / 6784.0 + 92097.0 / 339200.0, 11.0 / 84.0 - 187.0 / 2100.0, -1.0 / 40.0}}}},
    /* NSRK78::step (NSRK78.cu:22-116); its ERROR_K weights as passed there */
    {"NSRK78", 13, 1.0 / 8.0, 26, 27,
     {K1_,
      {'I', 1, 0, 1, {1.0 / 1000.0}},
      KB(2, 0.001),
      {'I', 2, 0, 2, {-491.0 / 81.0, 500.0 / 81.0}},
      KB(3, 0.1111111111),
      {'I', 3, 0, 3, {1.0 / 24.0, 0.0, 1.0 / 8.0}},
      KB(4, 0.1666666667),
      {'I', 4, 0, 4, {5.0 / 12.0, 0.0, -25.0 / 16.0, 25.0 / 16.0}},
      KB(5, 0.4166666667),
      {'I', 5, 0, 5, {1.0 / 20.0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 5.0}},
      KB(6, 0.5),
      {'I', 6, 0, 6, {-43.0 / 180.0, 0.0, 0.0, 127.0 / 108.0, -329.0 / 135.0, 7.0 / 3.0}},
      KB(7, 0.8333333333),
      {'I', 7, 0, 7, {27931.0 / 240300.0, 0.0, 0.0, -631.0 / 16020.0, 2459.0 / 6675.0, -3572.0 / 12015.0, 5.0 / 267.0}},
      KB(8, 0.1666666667),
      {'I', 8, 0, 8, {26114.0 / 12015.0, 0.0, 0.0, -7480.0 / 801.0, 67264.0 / 4005.0, -30640.0 / 2403.0, 1051.0 / 1335.0, 3.0}},
      KB(9, 0.6666666667),
      {'I', 9, 0, 9, {33096587331.0 / 17382812500.0, 0.0, 0.0, -1828977848.0 / 173828125.0, 62801809904.0 / 4345703125.0, -9389764774.0 / 869140625.0, 6380757669.0 / 8691406250.0, 98417891.0 / 19531250.0, -1692691.0 / 39062500.0}},
      KB(10, 0.76),
      {'I', 10, 0, 10, {-1456295425347.0 / 2642187500000.0, 0.0, 0.0, 110740056.0 / 34765625.0, -21221682384.0 / 4345703125.0, 58859060169.0 / 13906250000.0, -177381525069.0 / 1529687500000.0, -28942485159.0 / 27812500000.0, -1272297.0 / 312500000.0, 5151.0 / 297616.0}},
      KB(11, 0.84),
      {'I', 11, 0, 11, {844300798.0 / 137013275.0, 0.0, 0.0, -9997568.0 / 206035.0, 49636624.0 / 1030175.0, -3358834871.0 / 91067470.0, -40456983.0 / 1813108.0, 495817135.0 / 16647628.0, -149375.0 / 84266.0, 7470703125.0 / 1567431866.0, 1562500000.0 / 72342361.0}},
      KB(12, 1.0),
      {'I', 12, 0, 12, {-26225423.0 / 37371100.0, 0.0, 0.0, 807744.0 / 98345.0, -205584.0 / 37825.0, 5882202.0 / 1278485.0, -84543.0 / 432718.0, -223415.0 / 39338.0, -3625.0 / 6188.0, 292968750.0 / 374084711.0, 0.0, 0.0}},
      KB(13, 1.0),
      {'F', 13, 0, 13, {4241.0 / 88200.0, 0.0, 0.0, 0.0, 0.0, 9946.0 / 23205.0, -10449.0 / 1925.0, 415449.0 / 1573075.0, -2025.0 / 5096.0, 48828125.0 / 44900856.0, 48828125.0 / 9843561.0, 463.0 / 12600.0, 0.0}}},
     {K1_,
      {'I', 1, 0, 1, {1.0 / 1000.0}},
      KB(2, 0.001),
      {'I', 2, 0, 2, {-491.0 / 81.0, 500.0 / 81.0}},
      KB(3, 0.1111111111),
      {'I', 3, 0, 3, {1.0 / 24.0, 0.0, 1.0 / 8.0}},
      KB(4, 0.1666666667),
      {'I', 4, 0, 4, {5.0 / 12.0, 0.0, -25.0 / 16.0, 25.0 / 16.0}},
      KB(5, 0.4166666667),
      {'I', 5, 0, 5, {1.0 / 20.0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 5.0}},
      KB(6, 0.5),
      {'I', 6, 0, 6, {-43.0 / 180.0, 0.0, 0.0, 127.0 / 108.0, -329.0 / 135.0, 7.0 / 3.0}},
      KB(7, 0.8333333333),
      {'I', 7, 0, 7, {27931.0 / 240300.0, 0.0, 0.0, -631.0 / 16020.0, 2459.0 / 6675.0, -3572.0 / 12015.0, 5.0 / 267.0}},
      KB(8, 0.1666666667),
      {'I', 8, 0, 8, {26114.0 / 12015.0, 0.0, 0.0, -7480.0 / 801.0, 67264.0 / 4005.0, -30640.0 / 2403.0, 1051.0 / 1335.0, 3.0}},
      KB(9, 0.6666666667),
      {'I', 9, 0, 9, {33096587331.0 / 17382812500.0, 0.0, 0.0, -1828977848.0 / 173828125.0, 62801809904.0 / 4345703125.0, -9389764774.0 / 869140625.0, 6380757669.0 / 8691406250.0, 98417891.0 / 19531250.0, -1692691.0 / 39062500.0}},
      KB(10, 0.76),
      {'I', 10, 0, 10, {-1456295425347.0 / 2642187500000.0, 0.0, 0.0, 110740056.0 / 34765625.0, -21221682384.0 / 4345703125.0, 58859060169.0 / 13906250000.0, -177381525069.0 / 1529687500000.0, -28942485159.0 / 27812500000.0, -1272297.0 / 312500000.0, 5151.0 / 297616.0}},
      KB(11, 0.84),
      {'I', 11, 0, 11, {844300798.0 / 137013275.0, 0.0, 0.0, -9997568.0 / 206035.0, 49636624.0 / 1030175.0, -3358834871.0 / 91067470.0, -40456983.0 / 1813108.0, 495817135.0 / 16647628.0, -149375.0 / 84266.0, 7470703125.0 / 1567431866.0, 1562500000.0 / 72342361.0}},
      KB(12, 1.0),
      {'I', 12, 0, 12, {-26225423.0 / 37371100.0, 0.0, 0.0, 807744.0 / 98345.0, -205584.0 / 37825.0, 5882202.0 / 1278485.0, -84543.0 / 432718.0, -223415.0 / 39338.0, -3625.0 / 6188.0, 292968750.0 / 374084711.0, 0.0, 0.0}},
      KB(13, 1.0),
      {'I', 13, 0, 13, {4241.0 / 88200.0, 0.0, 0.0, 0.0, 0.0, 9946.0 / 23205.0, -10449.0 / 1925.0, 415449.0 / 1573075.0, -2025.0 / 5096.0, 48828125.0 / 44900856.0, 48828125.0 / 9843561.0, 463.0 / 12600.0, 0.0}},
      {'E', 13, 0, 13, {4241.0 / 88200.0 - 3799.0 / 79800.0, 0.0, 0.0, 0.0, 0.0, 9946.0 / 23205.0 - 538.0 / 1365.0, -10449.0 / 1925.0 - 351.0 / 1925.0, 415449.0 / 1573075.0 - 4149.0 / 15575.0, -2025.0 / 5096.0 - (-45.0 / 392.0), 48828125.0 / 44900856.0 - 48828125.0 / 284372088.0, 48828125.0 / 9843561.0 - 0.0, 463.0 / 12600.0 - 0.0, 0.0 - 221.0 / 4200.0}}}},
    /* DormandPrince85 coefficients (dormand_prince.cu:128-205); the reference swaps its fixed and adaptive branches: here the fixed step applies b and the adaptive step estimates its error with b - b5, b5 the weights its ERROR_K call passes */
    {"DP85", 12, 1.0 / 4.0, 24, 25,
     {K1_,
      {'I', 1, 0, 1, {5.260015195876773e-2}},
      KB(2, 0.05260015196),
      {'I', 2, 0, 2, {1.9725056984537899e-2, 5.9175170953613698e-2}},
      KB(3, 0.07890022794),
      {'I', 3, 0, 3, {2.9587585476806849e-2, 0.0, 8.8762756430420548e-2}},
      KB(4, 0.1183503419),
      {'I', 4, 0, 4, {2.4136513415926669e-1, 0.0, -8.8454947932828610e-1, 9.2483400326179200e-1}},
      KB(5, 0.2816496581),
      {'I', 5, 0, 5, {3.7037037037037037e-2, 0.0, 0.0, 1.7082860872947387e-1, 1.2546768756682243e-1}},
      KB(6, 0.3333333333),
      {'I', 6, 0, 6, {3.7109375e-2, 0.0, 0.0, 1.7025221101954404e-1, 6.0216538980455961e-2, -1.7578125e-2}},
      KB(7, 0.25),
      {'I', 7, 0, 7, {3.7092000118504793e-2, 0.0, 0.0, 1.7038392571223999e-1, 1.0726203044637328e-1, -1.5319437748624402e-2, 8.2737891638140229e-3}},
      KB(8, 0.3076923077),
      {'I', 8, 0, 8, {6.2411095871607572e-1, 0.0, 0.0, -3.3608926294469413e+0, -8.6821934684172601e-1, 2.7592099699446708e+1, 2.0154067550477893e+1, -4.3489884181069959e+1}},
      KB(9, 0.6512820513),
      {'I', 9, 0, 9, {4.7766253643826437e-1, 0.0, 0.0, -2.4881146199716676e+0, -5.9029082683684300e-1, 2.1230051448181194e+1, 1.5279233632882424e+1, -3.3288210968984863e+1, -2.0331201708508626e-2}},
      KB(10, 0.6),
      {'I', 10, 0, 10, {-9.3714243008598733e-1, 0.0, 0.0, 5.1863724288440637e+0, 1.0914373489967296e+0, -8.1497870107469261e+0, -1.8520065659996960e+1, 2.2739487099350504e+1, 2.4936055526796524e+0, -3.0467644718982195e+0}},
      KB(11, 0.8571428571),
      {'I', 11, 0, 11, {2.2733101475165382e+0, 0.0, 0.0, -1.0534495466737250e+1, -2.0008720582248625e+0, -1.7958931863118799e+1, 2.7948884529419960e+1, -2.8589982771350237e+0, -8.8728569335306295e+0, 1.2360567175794303e+1, 6.4339274601576353e-1}},
      KB(12, 1.0),
      {'F', 12, 0, 12, {5.4293734116568762e-2, 0.0, 0.0, 0.0, 0.0, 4.4503128927524089e+0, 1.8915178993145004e+0, -5.8012039600105848e+0, 3.1116436695781989e-1, -1.5216094966251608e-1, 2.0136540080403035e-1, 4.4710615727772591e-2}}},
     {K1_,
      {'I', 1, 0, 1, {5.260015195876773e-2}},
      KB(2, 0.05260015196),
      {'I', 2, 0, 2, {1.9725056984537899e-2, 5.9175170953613698e-2}},
      KB(3, 0.07890022794),
      {'I', 3, 0, 3, {2.9587585476806849e-2, 0.0, 8.8762756430420548e-2}},
      KB(4, 0.1183503419),
      {'I', 4, 0, 4, {2.4136513415926669e-1, 0.0, -8.8454947932828610e-1, 9.2483400326179200e-1}},
      KB(5, 0.2816496581),
      {'I', 5, 0, 5, {3.7037037037037037e-2, 0.0, 0.0, 1.7082860872947387e-1, 1.2546768756682243e-1}},
      KB(6, 0.3333333333),
      {'I', 6, 0, 6, {3.7109375e-2, 0.0, 0.0, 1.7025221101954404e-1, 6.0216538980455961e-2, -1.7578125e-2}},
      KB(7, 0.25),
      {'I', 7, 0, 7, {3.7092000118504793e-2, 0.0, 0.0, 1.7038392571223999e-1, 1.0726203044637328e-1, -1.5319437748624402e-2, 8.2737891638140229e-3}},
      KB(8, 0.3076923077),
      {'I', 8, 0, 8, {6.2411095871607572e-1, 0.0, 0.0, -3.3608926294469413e+0, -8.6821934684172601e-1, 2.7592099699446708e+1, 2.0154067550477893e+1, -4.3489884181069959e+1}},
      KB(9, 0.6512820513),
      {'I', 9, 0, 9, {4.7766253643826437e-1, 0.0, 0.0, -2.4881146199716676e+0, -5.9029082683684300e-1, 2.1230051448181194e+1, 1.5279233632882424e+1, -3.3288210968984863e+1, -2.0331201708508626e-2}},
      KB(10, 0.6),
      {'I', 10, 0, 10, {-9.3714243008598733e-1, 0.0, 0.0, 5.1863724288440637e+0, 1.0914373489967296e+0, -8.1497870107469261e+0, -1.8520065659996960e+1, 2.2739487099350504e+1, 2.4936055526796524e+0, -3.0467644718982195e+0}},
      KB(11, 0.8571428571),
      {'I', 11, 0, 11, {2.2733101475165382e+0, 0.0, 0.0, -1.0534495466737250e+1, -2.0008720582248625e+0, -1.7958931863118799e+1, 2.7948884529419960e+1, -2.8589982771350237e+0, -8.8728569335306295e+0, 1.2360567175794303e+1, 6.4339274601576353e-1}},
      KB(12, 1.0),
      {'I', 12, 0, 12, {5.4293734116568762e-2, 0.0, 0.0, 0.0, 0.0, 4.4503128927524089e+0, 1.8915178993145004e+0, -5.8012039600105848e+0, 3.1116436695781989e-1, -1.5216094966251608e-1, 2.0136540080403035e-1, 4.4710615727772591e-2}},
      {'E', 12, 0, 12, {5.4293734116568762e-2 - (4.1173689122373888e-2), 0.0, 0.0, 0.0, 0.0, 4.4503128927524089e+0 - (5.6754693391286133e+0), 1.8915178993145004e+0 - (2.3872768489717506e+0), -5.8012039600105848e+0 - (-7.4655811424655713e+0), 3.1116436695781989e-1 - (6.6149321570779360e-1), -1.5216094966251608e-1 - (-4.8634006837553356e-1), 2.0136540080403035e-1 - (1.1944219431891464e-1), 4.4710615727772591e-2 - (6.7065923591658886e-2)}}}},
};
#undef K1_
#undef KB

static const orc_iter* find_iter(const char* key) {
    for (size_t k = 0; k < sizeof(ITERS) / sizeof(ITERS[0]); k++)
        if (strcmp(ITERS[k].key, key) == 0) return &ITERS[k];
    return NULL;
}

static int kmax_of(const orc_op* ops, int n) {
    int m = 1;
    for (int k = 0; k < n; k++)
        if (ops[k].op == 'K' && ops[k].index > m) m = ops[k].index;
    return m;
}

/* ERROR_K (runge_sum_to_error, kernel_summation.cuh:270-289) on interior cells, psi+
   counted twice with -tetm (cuda_macro.cuh:244); norm = sum |buffer psi+|^2 on
   interior cells (Solver::adaptTimeStep, solver.cu:73-74). */
static void error_sum(const ws_t* w, int nw, const double* wts, double* err_sum, double* norm) {
    const orc_params* p = w->p;
    int slot[13];
    for (int j = 0; j < nw; j++) slot[j] = j;
    double es = 0.0, ns = 0.0;
    for (int g = 0; g < w->S; g++)
        for (int r = w->h; r < w->h + w->sgr; r++)
            for (int c = w->h; c < w->h + w->sgc; c++) {
                size_t i = (size_t)g * w->N2h + (size_t)r * w->ro + c;
                double s[2];
                ksum(w, w->kpsi[0], 2, i, nw, slot, wts, s);
                double er = w->dt * s[0], ei = w->dt * s[1];
                double e = er * er + ei * ei;
                es += p->tetm ? 2.0 * e : e;
                ns += w->bpsi[0][2 * i] * w->bpsi[0][2 * i] + w->bpsi[0][2 * i + 1] * w->bpsi[0][2 * i + 1];
            }
    *err_sum = es;
    *norm = ns;
}

/* normalizeImaginaryTimePropagation (solver_normalize_imag_time.cu:13-45) on the interior cells:
   psi_k *= sqrt(amp / (S_k dV)), S_k = sum |psi_k|^2 (< 1e-10 -> 1).  The reference sums its padded
   subgrids including halo cells; this sums each grid cell once.  Without a reservoir the reservoir
   normalisation of the reference scales zeros. */
static void ws_imag_normalize_field(ws_t* w, double* const* f, int el);

static void ws_imag_normalize(ws_t* w) {
    ws_imag_normalize_field(w, w->psi, 2);
    if (w->p->use_reservoir) ws_imag_normalize_field(w, w->n, w->nel);
}

/* f[comp] *= sqrt(amp / (sum |f[comp]|^2 dV)) over the interior cells (el doubles per cell) */
static void ws_imag_normalize_field(ws_t* w, double* const* f, int el) {
    const orc_params* p = w->p;
    for (int comp = 0; comp < w->ncomp; comp++) {
        double sum = 0.0;
        for (int g = 0; g < w->S; g++)
            for (int r = w->h; r < w->h + w->sgr; r++)
                for (int c = w->h; c < w->h + w->sgc; c++) {
                    size_t i = (size_t)g * w->N2h + (size_t)r * w->ro + c;
                    for (int e = 0; e < el; e++) sum += f[comp][el * i + e] * f[comp][el * i + e];
                }
        if (sum < 1e-10) sum = 1.0;
        const double sc = sqrt(p->imag_time_amplitude / (sum * p->dV));
        for (int g = 0; g < w->S; g++)
            for (int r = w->h; r < w->h + w->sgr; r++)
                for (int c = w->h; c < w->h + w->sgc; c++) {
                    size_t i = (size_t)g * w->N2h + (size_t)r * w->ro + c;
                    for (int e = 0; e < el; e++) f[comp][el * i + e] *= sc;
                }
    }
}

void orc_imag_normalize(const orc_params* p, double* psi_plus, double* psi_minus) {
    const size_t N = (size_t)p->N_c * p->N_r;
    double* f[2] = {psi_plus, psi_minus};
    for (int comp = 0; comp < 2; comp++) {
        if (!f[comp]) continue;
        double sum = 0.0;
        for (size_t i = 0; i < N; i++) sum += f[comp][2 * i] * f[comp][2 * i] + f[comp][2 * i + 1] * f[comp][2 * i + 1];
        if (sum < 1e-10) sum = 1.0;
        const double s = sqrt(p->imag_time_amplitude / (sum * p->dV));
        for (size_t i = 0; i < 2 * N; i++) f[comp][i] *= s;
    }
}

void orc_imag_normalize_reservoir(const orc_params* p, double* n_plus, double* n_minus) {
    const size_t N = (size_t)p->N_c * p->N_r;
    const int el = p->complex_reservoir ? 2 : 1;
    double* f[2] = {n_plus, n_minus};
    for (int comp = 0; comp < 2; comp++) {
        if (!f[comp]) continue;
        double sum = 0.0;
        for (size_t i = 0; i < el * N; i++) sum += f[comp][i] * f[comp][i];
        if (sum < 1e-10) sum = 1.0;
        const double s = sqrt(p->imag_time_amplitude / (sum * p->dV));
        for (size_t i = 0; i < el * N; i++) f[comp][i] *= s;
    }
}

/* Run one step() sequence on the workspace; returns 1 if the result is in the buffer. */
static int run_ops(ws_t* w, const orc_op* ops, int n, double* err_sum, double* norm) {
    int in_buffer = 0;
    sync_state(w);
    for (int k = 0; k < n; k++) {
        const orc_op* o = &ops[k];
        switch (o->op) {
            case 'K': calc_k(w, o->index, o->w, o->index > 1); break;
            case 'I': sum_to_buffer(w, o->index, o->nw, o->wts); in_buffer = 1; break;
            case 'F': final_sum(w, o->index, o->nw, o->wts); in_buffer = 0; break;
            case 'E': error_sum(w, o->nw, o->wts, err_sum, norm); break;
        }
    }
    return in_buffer;
}

int orc_iter_info(const char* key, int* halo, double* power, int* stages_fixed, int* stages_adaptive) {
    const orc_iter* it = find_iter(key);
    if (!it) return -1;
    if (halo) *halo = it->halo;
    if (power) *power = it->power;
    if (stages_fixed) *stages_fixed = kmax_of(it->fixed, it->nfixed);
    if (stages_adaptive) *stages_adaptive = it->nadapt ? kmax_of(it->adapt, it->nadapt) : 0;
    return 0;
}

int orc_iter_run(const orc_params* p, const char* key, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                 const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                 double dt, int steps, int64_t step0) {
    const orc_iter* it = find_iter(key);
    if (!it) return -3;
    ws_t w;
    if (ws_init(&w, p, it->halo)) return -1;
    w.dt = dt;
    load_state(&w, psi_plus, psi_minus, n_plus, n_minus, pump_plus, pump_minus, pot_plus, pot_minus);
    for (int s = 0; s < steps; s++) {
        gen_noise(&w, step0 + s, dt);
        double es, ns;
        run_ops(&w, it->fixed, it->nfixed, &es, &ns);
        if (p->imag_time_amplitude != 0.0) ws_imag_normalize(&w);
    }
    store_state(&w, psi_plus, psi_minus, n_plus, n_minus, 0);
    ws_free(&w);
    return 0;
}

int orc_iter_attempt(const orc_params* p, const char* key, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                     const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                     double dt, double noise_dt, int64_t step_index, double* err_sum, double* norm) {
    const orc_iter* it = find_iter(key);
    if (!it) return -3;
    if (!it->nadapt) return -2;
    ws_t w;
    if (ws_init(&w, p, it->halo)) return -1;
    w.dt = dt;
    load_state(&w, psi_plus, psi_minus, n_plus, n_minus, pump_plus, pump_minus, pot_plus, pot_minus);
    gen_noise(&w, step_index, noise_dt);
    *err_sum = 0.0;
    *norm = 0.0;
    run_ops(&w, it->adapt, it->nadapt, err_sum, norm);
    store_state(&w, psi_plus, psi_minus, n_plus, n_minus, 1); /* swapBuffers (solver.cu:107-114) */
    ws_free(&w);
    return 0;
}

int orc_adapt_dt_pow(double err_sum, double norm, double tolerance, double dt, double dt_min, double dt_max, double power,
                     double* new_dt, double* final_error) {
    double fe = fabs(err_sum / norm);
    double dh_arg = tolerance / 2.0 / (fe > DBL_MIN ? fe : DBL_MIN);
    double dh = pow(dh_arg, power);
    if (isnan(dh)) { dh = 0.9; fe = DBL_MAX; }
    if (isnan(fe)) { dh = 0.9; fe = DBL_MAX; }
    double nd = dt * dh < dt_max ? dt * dh : dt_max;
    if (nd < dt_min) nd = dt_min;
    *new_dt = nd;
    *final_error = fe;
    return fe < tolerance;
}

int64_t orc_iter_adaptive_run(const orc_params* p, const char* key, double* psi_plus, double* psi_minus, double* n_plus,
                              double* n_minus, const double* pump_plus, const double* pump_minus, const double* pot_plus,
                              const double* pot_minus, double* t, double* dt, double t_max, double tolerance, double dt_min,
                              double dt_max, int max_rejections, int64_t* n_rejected) {
    const orc_iter* it = find_iter(key);
    if (!it) return -3;
    if (!it->nadapt) return -2;
    const size_t N = (size_t)p->N_c * p->N_r;
    const int nc = p->tetm ? 2 : 1;
    double *ps[2] = {psi_plus, psi_minus}, *ns[2] = {n_plus, n_minus};
    double *tp[2] = {0, 0}, *tn[2] = {0, 0};
    for (int c = 0; c < nc; c++) {
        tp[c] = malloc(2 * N * sizeof(double));
        tn[c] = malloc((p->complex_reservoir ? 2 : 1) * N * sizeof(double));
    }
    int64_t accepted = 0, rejected = 0, step = 0;
    int64_t rc = 0;
    while (*t < t_max) {
        const double noise_dt = *dt; /* noise drawn once per iterate() with sqrt(dt) (solver.cu:36-45) */
        *t = *t + *dt;               /* solver.cu:50, before the attempts */
        int rej = 0;
        for (;;) {
            for (int c = 0; c < nc; c++) {
                memcpy(tp[c], ps[c], 2 * N * sizeof(double));
                if (p->use_reservoir) memcpy(tn[c], ns[c], (p->complex_reservoir ? 2 : 1) * N * sizeof(double));
            }
            double es, nm, nd, fe;
            if (orc_iter_attempt(p, key, tp[0], tp[1], tn[0], tn[1], pump_plus, pump_minus, pot_plus, pot_minus, *dt, noise_dt, step,
                                 &es, &nm)) {
                rc = -1;
                goto done;
            }
            int acc = orc_adapt_dt_pow(es, nm, tolerance, *dt, dt_min, dt_max, it->power, &nd, &fe);
            *dt = nd;
            if (acc) break;
            rejected++;
            if (++rej > max_rejections) {
                rc = -2;
                goto done;
            }
        }
        for (int c = 0; c < nc; c++) {
            memcpy(ps[c], tp[c], 2 * N * sizeof(double));
            if (p->use_reservoir) memcpy(ns[c], tn[c], (p->complex_reservoir ? 2 : 1) * N * sizeof(double));
        }
        if (p->imag_time_amplitude != 0.0) {
            orc_imag_normalize(p, ps[0], nc > 1 ? ps[1] : NULL);
            if (p->use_reservoir) orc_imag_normalize_reservoir(p, ns[0], nc > 1 ? ns[1] : NULL);
        }
        accepted++;
        step++;
    }
    rc = accepted;
done:
    for (int c = 0; c < nc; c++) { free(tp[c]); free(tn[c]); }
    *n_rejected = rejected;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* FFT filter (solver_fft.cu:67-119, kernel_fft.cuh:15-39)                  */
/* ------------------------------------------------------------------------ */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* 1D unnormalised DFT of length n with stride; radix-2 iterative for powers
   of two, direct summation otherwise. */
static void dft1d(double* x, int n, size_t stride, int sign, double* tmp) {
    for (int i = 0; i < n; i++) { tmp[2 * i] = x[2 * i * stride]; tmp[2 * i + 1] = x[2 * i * stride + 1]; }
    if (is_pow2(n)) {
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double a = tmp[2 * i], b = tmp[2 * i + 1];
                tmp[2 * i] = tmp[2 * j]; tmp[2 * i + 1] = tmp[2 * j + 1];
                tmp[2 * j] = a; tmp[2 * j + 1] = b;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < len / 2; k++) {
                    double ang = sign * 2.0 * M_PI * k / len;
                    double wr = cos(ang), wi = sin(ang);
                    double* u = tmp + 2 * (i + k);
                    double* v = tmp + 2 * (i + k + len / 2);
                    double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
                    v[0] = u[0] - vr; v[1] = u[1] - vi;
                    u[0] += vr; u[1] += vi;
                }
        }
        for (int i = 0; i < n; i++) { x[2 * i * stride] = tmp[2 * i]; x[2 * i * stride + 1] = tmp[2 * i + 1]; }
    } else {
        double* o = tmp + 2 * n;
        for (int k = 0; k < n; k++) {
            double sr = 0, si = 0;
            for (int j = 0; j < n; j++) {
                double ang = sign * 2.0 * M_PI * (double)((long)j * k % n) / n;
                sr += tmp[2 * j] * cos(ang) - tmp[2 * j + 1] * sin(ang);
                si += tmp[2 * j] * sin(ang) + tmp[2 * j + 1] * cos(ang);
            }
            o[2 * k] = sr; o[2 * k + 1] = si;
        }
        for (int i = 0; i < n; i++) { x[2 * i * stride] = o[2 * i]; x[2 * i * stride + 1] = o[2 * i + 1]; }
    }
}

int orc_fft2d(double* data, int N_c, int N_r, int sign) {
    int m = N_c > N_r ? N_c : N_r;
    double* tmp = malloc(4 * (size_t)m * sizeof(double));
    if (!tmp) return -1;
    for (int r = 0; r < N_r; r++) dft1d(data + 2 * (size_t)r * N_c, N_c, 1, sign, tmp);
    for (int c = 0; c < N_c; c++) dft1d(data + 2 * (size_t)c, N_r, (size_t)N_c, sign, tmp);
    free(tmp);
    return 0;
}

int orc_fft_filter(double* psi, const double* mask, int N_c, int N_r) {
    const size_t N = (size_t)N_c * N_r;
    if (orc_fft2d(psi, N_c, N_r, -1)) return -1;
    /* fftshift of the mask by quadrant swap (fft_shift_2D), then F <- F / N * mask */
    for (int r = 0; r < N_r; r++)
        for (int c = 0; c < N_c; c++) {
            int rs = (r + N_r / 2) % N_r, cs = (c + N_c / 2) % N_c;
            double m = mask[(size_t)rs * N_c + cs];
            double* f = psi + 2 * ((size_t)r * N_c + c);
            f[0] = f[0] / (double)N * m;
            f[1] = f[1] / (double)N * m;
        }
    return orc_fft2d(psi, N_c, N_r, +1);
}

/* ------------------------------------------------------------------------ */
/* Split-step Fourier (split_step_fourier_method.cu:25-71)                  */
/* ------------------------------------------------------------------------ */
/* gp_scalar_linear_fourier (kernel_gp_compute.cuh:244-262) between a forward and an inverse
   transform: F <- F / N2 * exp(-i hbar/(2 m) (kx^2 + ky^2) dt / 2) */
static int ssfm_linear(const orc_params* p, double* f, double dt) {
    const int Nc = p->N_c, Nr = p->N_r;
    const double N2 = (double)Nc * Nr;
    if (orc_fft2d(f, Nc, Nr, -1)) return -1;
    for (int r = 0; r < Nr; r++)
        for (int c = 0; c < Nc; c++) {
            const double kx = 2.0 * 3.1415926535 * (double)(c <= Nc / 2 ? c : c - Nc) / p->L_x;
            const double ky = 2.0 * 3.1415926535 * (double)(r <= Nr / 2 ? r : r - Nr) / p->L_y;
            const double lin = p->h_bar_s / 2.0 / p->m_eff * (kx * kx + ky * ky);
            const double ph = -lin * dt / 2.0;
            double* v = f + 2 * ((size_t)r * Nc + c);
            const double ar = v[0] / N2, ai = v[1] / N2;
            v[0] = ar * cos(ph) - ai * sin(ph);
            v[1] = ar * sin(ph) + ai * cos(ph);
        }
    return orc_fft2d(f, Nc, Nr, +1);
}

int orc_ssfm_run(const orc_params* p, double* psi_plus, double* psi_minus, double* n_plus, double* n_minus,
                 const double* pump_plus, const double* pump_minus, const double* pot_plus, const double* pot_minus,
                 double dt, int steps, int64_t step0) {
    const size_t N = (size_t)p->N_c * p->N_r;
    const int nc = p->tetm ? 2 : 1;
    /* complex reservoir: n held as {re, im} (a complex temporal pump amplitude enters it whole) */
    const int cres = p->complex_reservoir && p->use_reservoir, nel = cres ? 2 : 1;
    double* psi[2] = {psi_plus, psi_minus};
    double* n[2] = {n_plus, n_minus};
    const double* pump[2] = {pump_plus, pump_minus};
    const double* pot[2] = {pot_plus, pot_minus};
    const int stoch = p->use_stochastic && p->stochastic_amplitude > 0.0;
    double* nn[2] = {0, 0};
    double* xi = calloc(2 * N, sizeof(double));
    for (int k = 0; k < nc; k++) nn[k] = calloc(nel * N, sizeof(double));
    for (int st = 0; st < steps; st++) {
        /* per-step noise: amplitude * sqrt(dt) * N2(seed, step, cell) (solver.cu:36-45) */
        if (stoch)
            for (size_t i = 0; i < N; i++) {
                double re, im;
                orc_noise_normal2(p->noise_seed, (uint64_t)(step0 + st), (uint32_t)i, 0, &re, &im);
                xi[2 * i] = p->stochastic_amplitude * sqrt(dt) * re;
                xi[2 * i + 1] = p->stochastic_amplitude * sqrt(dt) * im;
            }
        for (int k = 0; k < nc; k++)
            if (ssfm_linear(p, psi[k], dt)) return -1;
        /* gp_scalar_nonlinear (kernel_gp_compute.cuh:264-382) */
        for (size_t i = 0; i < N; i++) {
            double nrm[2] = {0, 0}, nv[2] = {0, 0}, nvi[2] = {0, 0};
            for (int k = 0; k < nc; k++) {
                nrm[k] = psi[k][2 * i] * psi[k][2 * i] + psi[k][2 * i + 1] * psi[k][2 * i + 1];
                nv[k] = p->use_reservoir ? n[k][nel * i] : 0.0;
                nvi[k] = cres ? n[k][2 * i + 1] : 0.0;
            }
            for (int k = 0; k < nc; k++) {
                /* result = (g_c |psi|^2, -hbar/2 gamma_c) + V a + g_r n + i hbar/2 R n (+ g_pm |psi_o|^2) (- g_c/dV) */
                double A = p->g_c * nrm[k], B = -p->h_bar_s * 0.5 * p->gamma_c;
                if (p->src_groups[1] > 0) {
                    /* potential chain: result += V_g a_g with the amplitude of the step start (:276-279) */
                    for (int g = 0; g < p->src_groups[1]; g++) {
                        const double v = p->src_field[1][g][k][i];
                        A += v * p->src_amp[1][g][0];
                        B += v * p->src_amp[1][g][1];
                    }
                } else if (p->use_potential) {
                    A += pot[k][i] * p->pot_amp;
                }
                A += p->g_r * nv[k];
                if (p->tetm) A += p->g_pm * nrm[1 - k];
                if (stoch) A -= p->g_c / p->dV;
                B += p->h_bar_s * 0.5 * p->R * nv[k];
                if (cres) {
                    /* Im n in g_r n + i hbar/2 R n: A -= hbar/2 R Im n, B += g_r Im n */
                    A -= p->h_bar_s * 0.5 * p->R * nvi[k];
                    B += p->g_r * nvi[k];
                }
                /* psi * exp((B - i A) dt / hbar) */
                const double er = B * p->one_over_h_bar_s * dt, ei = -A * p->one_over_h_bar_s * dt;
                const double m = exp(er), fr = m * cos(ei), fi = m * sin(ei);
                const double pr = psi[k][2 * i], pi = psi[k][2 * i + 1];
                psi[k][2 * i] = pr * fr - pi * fi;
                psi[k][2 * i + 1] = pr * fi + pi * fr;
                if (p->use_reservoir) {
                    double res = -p->gamma_r * nv[k], resi = -p->gamma_r * nvi[k];
                    res -= p->R * nrm[k] * nv[k];
                    resi -= p->R * nrm[k] * nvi[k];
                    if (p->src_groups[0] > 0) {
                        for (int g = 0; g < p->src_groups[0]; g++) {
                            res += p->src_field[0][g][k][i] * p->src_amp[0][g][0];
                            resi += p->src_field[0][g][k][i] * p->src_amp[0][g][1];
                        }
                    } else if (p->use_pump) {
                        res += pump[k][i] * p->pump_amp;
                    }
                    if (stoch) {
                        res += p->R * nv[k] / p->dV;
                        resi += p->R * nvi[k] / p->dV;
                    }
                    nn[k][nel * i] = nv[k] + res * dt;
                    if (cres) nn[k][2 * i + 1] = nvi[k] + resi * dt;
                }
            }
        }
        for (int k = 0; k < nc; k++)
            if (ssfm_linear(p, psi[k], dt)) return -1;
        /* gp_scalar_independent (kernel_gp_compute.cuh:387-440) */
        for (size_t i = 0; i < N; i++) {
            for (int k = 0; k < nc; k++) {
                double rr = 0.0, ri = 0.0;
                /* pulse chain: scalar (dt/hbar) F_g a_g (:392-396), TE-TM (-i dt/hbar) F_g a_g (:409-413) */
                for (int g = 0; g < p->src_groups[2]; g++) {
                    const double* F = p->src_field[2][g][k];
                    const double ar = p->src_amp[2][g][0], ai = p->src_amp[2][g][1];
                    const double zr = p->one_over_h_bar_s * dt * (F[2 * i] * ar - F[2 * i + 1] * ai);
                    const double zi = p->one_over_h_bar_s * dt * (F[2 * i] * ai + F[2 * i + 1] * ar);
                    rr += p->tetm ? zi : zr;
                    ri += p->tetm ? -zr : zi;
                }
                if (stoch && cres) {
                    /* complex reservoir: principal square root, complex product with xi */
                    const double* nsrc = p->tetm ? nn[k] : n[0];
                    const double complex z = ((p->R * nsrc[2 * i] + p->gamma_c) + I * (p->R * nsrc[2 * i + 1])) / (4.0 * p->dV);
                    const double complex sq = csqrt(z);
                    rr += xi[2 * i] * creal(sq) - xi[2 * i + 1] * cimag(sq);
                    ri += xi[2 * i] * cimag(sq) + xi[2 * i + 1] * creal(sq);
                } else if (stoch) {
                    /* scalar: the reservoir of the step start; TE-TM: the new one of the component */
                    const double nr = p->use_reservoir ? (p->tetm ? nn[k][i] : n[0][i]) : 0.0;
                    const double sg = sqrt((p->R * nr + p->gamma_c) / (4.0 * p->dV));
                    rr += xi[2 * i] * sg;
                    ri += xi[2 * i + 1] * sg;
                }
                psi[k][2 * i] += rr;
                psi[k][2 * i + 1] += ri;
            }
            if (p->use_reservoir && !p->tetm)
                for (int e = 0; e < nel; e++) n[0][nel * i + e] = nn[0][nel * i + e];
        }
        /* normalizeImaginaryTimePropagation after every iterate (solver.cu:55-56): the split-step step
           itself takes no -i dt (split_step_fourier_method.cu:25-71), so psi and the real reservoir are
           only rescaled */
        if (p->imag_time_amplitude != 0.0) {
            orc_imag_normalize(p, psi[0], nc > 1 ? psi[1] : NULL);
            if (p->use_reservoir) orc_imag_normalize_reservoir(p, n[0], nc > 1 ? n[1] : NULL);
        }
    }
    free(xi);
    for (int k = 0; k < nc; k++) free(nn[k]);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Envelope (envelope.cpp:332-412)                                          */
/* ------------------------------------------------------------------------ */
void orc_envelope_add(double* out, int N_c, int N_r, double L_x, double L_y, double dx, double dy,
                      double amp, double width_x, double width_y, double pos_x, double pos_y,
                      double exponent, int type_flags, double charge_m, double k0_x, double k0_y, int behaviour) {
    for (int row = 0; row < N_r; row++)
        for (int col = 0; col < N_c; col++) {
            size_t i = (size_t)row * N_c + col;
            double cx = -L_x / 2.0 + dx * col;
            double cy = -L_y / 2.0 + dy * row;
            if (type_flags & 16) { /* Local */
                cx = -1.0 + 2.0 * col / (N_c - 1);
                cy = -1.0 + 2.0 * row / (N_r - 1);
            }
            double ar = amp, ai = 0.0;
            double ex = 0.5 * (((cx - pos_x) / width_x) * ((cx - pos_x) / width_x) + ((cy - pos_y) / width_y) * ((cy - pos_y) / width_y));
            double ef = exp(-pow(ex, exponent));
            if (type_flags & 2) ef = pow(exp(-ex), exponent); /* OuterExponent */
            double pre = (type_flags & 4) ? ex : 1.0;         /* Ring */
            double phi = charge_m * atan2(cx - pos_x, cy - pos_y);
            double chr = cos(phi), chi = sin(phi);
            if (!(type_flags & 8)) { /* not NoDivide: A / sqrt(2*3.1415*wx*wy) */
                double d = sqrt(2 * 3.1415 * width_x * width_y);
                ar /= d; ai /= d;
            }
            double mp = k0_x * (cx - pos_x) + k0_y * (cy - pos_y);
            double mr = cos(mp), mi = sin(mp);
            if (behaviour & 8) { /* Adaptive: amplitude = amp * buffer */
                ar = amp * out[2 * i]; ai = amp * out[2 * i + 1];
            }
            if (behaviour & 16) { /* Complex: amplitude = i * real(amplitude) */
                ai = ar; ar = 0.0;
            }
            /* contribution = amplitude * pre * ef * charge * momentum */
            double s = pre * ef;
            double t1r = ar * s, t1i = ai * s;
            double t2r = t1r * chr - t1i * chi, t2i = t1r * chi + t1i * chr;
            double cr = t2r * mr - t2i * mi, ci = t2r * mi + t2i * mr;
            if (behaviour & 1) { out[2 * i] += cr; out[2 * i + 1] += ci; }
            else if (behaviour == 2) {
                double br = out[2 * i], bi = out[2 * i + 1];
                out[2 * i] = br * cr - bi * ci; out[2 * i + 1] = br * ci + bi * cr;
            } else if (behaviour == 4) { out[2 * i] = cr; out[2 * i + 1] = ci; }
        }
}

/* ======================================================================
 * Discrete adjoint of fixed-step RK4 (scalar model), FP64.
 *
 * Not a reference function (the reference has no adjoint): the gradient of
 * the forward map that orc_rk4_run restates (RungeKutta4::step,
 * runge_kutta.cu:74-94, RHS kernel_gp_compute.cuh:50-110), for checking the
 * HIP adjoint.  Written in complex form on a grid padded by 4 ghost cells in
 * every zero-boundary direction: the reference evaluates stage s on the halo
 * of width 4 - s, so stage states exist on ghost cells (from a zero input
 * halo) and feed the boundary cells; a padded global evaluation reproduces
 * exactly that map.  Periodic directions wrap instead.
 *
 * RHS (complex psi, real n):
 *   k_psi = -(i/hbar)[m~ L psi + (g_c|psi|^2 + g_r n + V~) psi] + (R n - gamma_c)/2 psi
 *   k_n   = -(gamma_r + R|psi|^2) n + P~
 * Cotangents use the real inner product <a, b> = Re(conj(a) b); for a
 * cotangent (mu, nu) of (k_psi, k_n) the pullback is
 *   psi_bar = (i/hbar)[m~ L mu + (g_c|psi|^2 + g_r n + V~) mu] + (i/hbar) g_c |psi|^2 mu
 *             - (i/hbar) g_c psi^2 conj(mu) + (R n - gamma_c)/2 mu - 2 R n nu psi
 *   n_bar   = Re(conj(mu) (R/2 - i g_r/hbar) psi) - (gamma_r + R|psi|^2) nu
 *   V_bar  += vamp Re(conj(mu) (-i/hbar) psi),   P_bar += pamp nu
 * (L is symmetric under both boundary conditions, so it is its own adjoint).
 * ====================================================================== */
#include <complex.h>

typedef struct {
    int W, H;    /* padded extents */
    int gx, gy;  /* ghost widths (0 for periodic directions) */
    int px, py;
} adj_grid_t;

static long adj_nb(const adj_grid_t* g, int r, int c) {
    if (g->py) r = ((r % g->H) + g->H) % g->H;
    else if (r < 0 || r >= g->H) return -1;
    if (g->px) c = ((c % g->W) + g->W) % g->W;
    else if (c < 0 || c >= g->W) return -1;
    return (long)r * g->W + c;
}

/* m~ L f on the padded grid (zero beyond it) */
static void adj_lap(const adj_grid_t* g, const orc_params* p, const double complex* f, double complex* out) {
    for (int r = 0; r < g->H; r++)
        for (int c = 0; c < g->W; c++) {
            long n = adj_nb(g, r - 1, c), s = adj_nb(g, r + 1, c), w = adj_nb(g, r, c - 1), e = adj_nb(g, r, c + 1);
            double complex ns = (n >= 0 ? f[n] : 0) + (s >= 0 ? f[s] : 0);
            double complex ew = (w >= 0 ? f[w] : 0) + (e >= 0 ? f[e] : 0);
            long i = (long)r * g->W + c;
            out[i] = p->m_eff_scaled * (p->m2_over_dx2_p_dy2 * f[i] + ns * p->one_over_dy2 + ew * p->one_over_dx2);
        }
}

static void adj_rhs(const adj_grid_t* g, const orc_params* p, const double complex* psi, const double* n, const double* P,
                    const double* V, double pamp, double vamp, double complex* kpsi, double* kn, double complex* tmp) {
    const long M = (long)g->W * g->H;
    const double ohb = p->one_over_h_bar_s;
    adj_lap(g, p, psi, tmp);
    for (long i = 0; i < M; i++) {
        const double nv = p->use_reservoir ? n[i] : 0.0;
        const double a2 = creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
        const double pot = (p->use_potential ? V[i] * vamp : 0.0) + p->g_c * a2 + p->g_r * nv;
        kpsi[i] = -I * ohb * (tmp[i] + pot * psi[i]) + 0.5 * (p->R * nv - p->gamma_c) * psi[i];
        if (p->use_reservoir) kn[i] = -(p->gamma_r + p->R * a2) * nv + (p->use_pump ? P[i] * pamp : 0.0);
    }
}

/* pullback of (mu, nu) through the RHS at (psi, n); accumulates into (opsi, on), gP, gV */
static void adj_vjp(const adj_grid_t* g, const orc_params* p, const double complex* psi, const double* n, const double* V,
                    double pamp, double vamp, const double complex* mu, const double* nu, double complex* opsi, double* on,
                    double* gP, double* gV, double complex* tmp) {
    const long M = (long)g->W * g->H;
    const double ohb = p->one_over_h_bar_s;
    adj_lap(g, p, mu, tmp);
    for (long i = 0; i < M; i++) {
        const double nv = p->use_reservoir ? n[i] : 0.0;
        const double nuv = p->use_reservoir ? nu[i] : 0.0;
        const double complex z = psi[i];
        const double a2 = creal(z) * creal(z) + cimag(z) * cimag(z);
        const double pot = (p->use_potential ? V[i] * vamp : 0.0) + p->g_c * a2 + p->g_r * nv;
        opsi[i] += I * ohb * (tmp[i] + pot * mu[i]) + I * ohb * p->g_c * a2 * mu[i] - I * ohb * p->g_c * z * z * conj(mu[i]) +
                   0.5 * (p->R * nv - p->gamma_c) * mu[i] - 2.0 * p->R * nv * nuv * z;
        if (p->use_reservoir) on[i] += creal(conj(mu[i]) * (0.5 * p->R - I * ohb * p->g_r) * z) - (p->gamma_r + p->R * a2) * nuv;
        if (gV) gV[i] += vamp * creal(conj(mu[i]) * (-I * ohb) * z);
        if (gP && p->use_reservoir) gP[i] += pamp * nuv;
    }
}

int orc_rk4_adjoint(const orc_params* p, const double* psi0, const double* n0, const double* pump, const double* pot, double dt,
                    int steps, double* lam_psi, double* lam_n, double* g_pump, double* g_pot, double* psi_T, double* n_T) {
    if (p->tetm || (p->use_stochastic && p->stochastic_amplitude > 0.0)) return -2;
    adj_grid_t g;
    g.gx = p->periodic_x ? 0 : 4;
    g.gy = p->periodic_y ? 0 : 4;
    g.W = p->N_c + 2 * g.gx;
    g.H = p->N_r + 2 * g.gy;
    g.px = p->periodic_x;
    g.py = p->periodic_y;
    const long M = (long)g.W * g.H;
    const long NT = (long)p->N_c * p->N_r;
    const double w[4] = {0.0, 0.5, 0.5, 1.0};
    double pamp[4], vamp[4];
    for (int s = 0; s < 4; s++) {
        pamp[s] = p->pump_amp * (1.0 - w[s]) + w[s] * p->pump_amp_next;
        vamp[s] = p->pot_amp * (1.0 - w[s]) + w[s] * p->pot_amp_next;
    }
    /* padded fields; states x_0 .. x_steps */
    double complex* X = xcalloc((size_t)(steps + 1) * M, sizeof(double complex));
    double* Xn = xcalloc((size_t)(steps + 1) * M, sizeof(double));
    double *P = xcalloc(M, sizeof(double)), *V = xcalloc(M, sizeof(double));
    double complex *Y[4], *K[4], *tmp = xcalloc(M, sizeof(double complex));
    double *Yn[4], *Kn[4];
    for (int s = 0; s < 4; s++) {
        Y[s] = xcalloc(M, sizeof(double complex));
        K[s] = xcalloc(M, sizeof(double complex));
        Yn[s] = xcalloc(M, sizeof(double));
        Kn[s] = xcalloc(M, sizeof(double));
    }
    double complex *lam = xcalloc(M, sizeof(double complex)), *cs = xcalloc(M, sizeof(double complex));
    double complex* yb = xcalloc(M, sizeof(double complex));
    double *lamn = xcalloc(M, sizeof(double)), *csn = xcalloc(M, sizeof(double)), *ybn = xcalloc(M, sizeof(double));
    double complex* acc = xcalloc(M, sizeof(double complex));
    double* accn = xcalloc(M, sizeof(double));
    double *gPp = xcalloc(M, sizeof(double)), *gVp = xcalloc(M, sizeof(double));
#define PAD(rr, cc) ((long)((rr) + g.gy) * g.W + (cc) + g.gx)
    for (int r = 0; r < p->N_r; r++)
        for (int c = 0; c < p->N_c; c++) {
            long i = (long)r * p->N_c + c, q = PAD(r, c);
            X[q] = psi0[2 * i] + I * psi0[2 * i + 1];
            Xn[q] = (p->use_reservoir && n0) ? n0[i] : 0.0;
            P[q] = (p->use_pump && pump) ? pump[i] : 0.0;
            V[q] = (p->use_potential && pot) ? pot[i] : 0.0;
            lam[q] = lam_psi[2 * i] + I * lam_psi[2 * i + 1];
            lamn[q] = (p->use_reservoir && lam_n) ? lam_n[i] : 0.0;
        }
    const double cY[4] = {0.0, 0.5 * dt, 0.5 * dt, dt};   /* Y_s = x + cY[s] k_{s-1} */
    const double bK[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
    /* stages of one step from state x (padded, ghosts zero) */
#define STAGES(x, xn)                                                                                     \
    for (int s = 0; s < 4; s++) {                                                                         \
        for (long i = 0; i < M; i++) {                                                                    \
            Y[s][i] = s ? (x)[i] + cY[s] * K[s - 1][i] : (x)[i];                                          \
            Yn[s][i] = s ? (xn)[i] + cY[s] * Kn[s - 1][i] : (xn)[i];                                      \
        }                                                                                                 \
        adj_rhs(&g, p, Y[s], Yn[s], P, V, pamp[s], vamp[s], K[s], Kn[s], tmp);                            \
    }
    for (int k = 0; k < steps; k++) {
        double complex* x = X + (size_t)k * M;
        double* xn = Xn + (size_t)k * M;
        STAGES(x, xn);
        double complex* x1 = X + (size_t)(k + 1) * M;
        double* x1n = Xn + (size_t)(k + 1) * M;
        for (int r = 0; r < p->N_r; r++)
            for (int c = 0; c < p->N_c; c++) {
                long q = PAD(r, c);
                x1[q] = x[q] + bK[0] * K[0][q] + bK[1] * K[1][q] + bK[2] * K[2][q] + bK[3] * K[3][q];
                x1n[q] = p->use_reservoir ? xn[q] + bK[0] * Kn[0][q] + bK[1] * Kn[1][q] + bK[2] * Kn[2][q] + bK[3] * Kn[3][q] : 0.0;
            }
    }
    /* reverse sweep */
    for (int k = steps - 1; k >= 0; k--) {
        double complex* x = X + (size_t)k * M;
        double* xn = Xn + (size_t)k * M;
        STAGES(x, xn);
        memset(acc, 0, M * sizeof(double complex));
        memset(accn, 0, M * sizeof(double));
        memset(yb, 0, M * sizeof(double complex));
        memset(ybn, 0, M * sizeof(double));
        for (int s = 3; s >= 0; s--) {
            /* c_s = b_s l' + cY[s+1] Yb_{s+1} */
            for (long i = 0; i < M; i++) {
                cs[i] = bK[s] * lam[i] + (s < 3 ? cY[s + 1] * yb[i] : 0.0);
                csn[i] = bK[s] * lamn[i] + (s < 3 ? cY[s + 1] * ybn[i] : 0.0);
            }
            memset(yb, 0, M * sizeof(double complex));
            memset(ybn, 0, M * sizeof(double));
            adj_vjp(&g, p, Y[s], Yn[s], V, pamp[s], vamp[s], cs, csn, yb, ybn, gPp, gVp, tmp);
            for (long i = 0; i < M; i++) {
                acc[i] += yb[i];
                accn[i] += ybn[i];
            }
        }
        /* l = l' + sum_s Yb_s on the grid; ghost cells are not variables (zero input halo) */
        for (long i = 0; i < M; i++) {
            lam[i] += acc[i];
            lamn[i] += accn[i];
        }
        for (int r = 0; r < g.H; r++)
            for (int c = 0; c < g.W; c++) {
                int rr = r - g.gy, cc = c - g.gx;
                if (rr < 0 || rr >= p->N_r || cc < 0 || cc >= p->N_c) {
                    lam[(long)r * g.W + c] = 0;
                    lamn[(long)r * g.W + c] = 0;
                }
            }
    }
    for (int r = 0; r < p->N_r; r++)
        for (int c = 0; c < p->N_c; c++) {
            long i = (long)r * p->N_c + c, q = PAD(r, c);
            lam_psi[2 * i] = creal(lam[q]);
            lam_psi[2 * i + 1] = cimag(lam[q]);
            if (lam_n) lam_n[i] = p->use_reservoir ? lamn[q] : 0.0;
            if (g_pump) g_pump[i] += gPp[q];
            if (g_pot) g_pot[i] += gVp[q];
            if (psi_T) {
                psi_T[2 * i] = creal(X[(size_t)steps * M + q]);
                psi_T[2 * i + 1] = cimag(X[(size_t)steps * M + q]);
            }
            if (n_T) n_T[i] = Xn[(size_t)steps * M + q];
        }
#undef PAD
#undef STAGES
    free(X); free(Xn); free(P); free(V); free(tmp);
    for (int s = 0; s < 4; s++) { free(Y[s]); free(K[s]); free(Yn[s]); free(Kn[s]); }
    free(lam); free(cs); free(yb); free(lamn); free(csn); free(ybn); free(acc); free(accn); free(gPp); free(gVp);
    (void)NT;
    return 0;
}
