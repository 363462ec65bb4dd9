/*
 * phx_tableaux.h -- Butcher tableaux of the explicit Runge-Kutta iterators and the
 * registry that maps the reference's --iterator keys onto them.
 *
 * Each tableau is the coefficient set the reference's step() passes to its
 * CALCULATE_K / INTERMEDIATE_SUM_K / FINAL_SUM_K / ERROR_K sequence (cited per
 * struct), including where that differs from the textbook method:
 *   - BoSha: the third stage input uses a32 = 3/9 (bogacki.cu:41), not 3/4;
 *   - CashKarp: a41 = 3/1 (cash_karp.cu:47), not 3/10; the error weight of K5 is
 *     -2187/6784 (cash_karp.cu:70), not -277/14336;
 *   - Fehlberg5: the error weight of K5 is -9/50 - 1/5 (fehlberg.cu:137), not -9/50 + 1/5;
 *   - SSPRK3: the third stage is evaluated at envelope weight 1 (SSPRK3.cu:38), not 1/2;
 *   - Ralston4: stage weights c3 = c4 = 0.51823725421 (ralston.cu:117,121), as passed.
 * Two reference sequences are undefined and are replaced by the tableau they state:
 *   - RK3 (runge_kutta.cu:46) sums a fourth K slot that the solver never writes
 *     (k_max_ = 3): here b = (1/6, 2/3, 1/6), as its own tableau string says;
 *   - Ralston4 (ralston.cu:92-123) evaluates four stages with k_max_ = halo = 3, so
 *     its fourth K lands outside the K buffer: here the four stages run with a
 *     halo of four, which is what the coefficients describe.
 *
 * Layout: S stages; a[s][j] (j < s) weights of stage s's input psi0 + dt sum_j a[s][j] K_j;
 * c[s] the envelope interpolation weight of stage s (gp_scalar's w); the output is
 * psi0 + dt sum_{j<NB} b[j] K_j; with ERR the embedded error is dt sum_j e[j] K_j(psi+)
 * over all S stages (FSAL methods evaluate stage S-1 on the output, a[S-1] = b).
 */
#pragma once

namespace phx {

enum class RkTableau {
    RK4,
    DP45_FIXED,
    DP45_ADAPTIVE,
    NEWTON,
    MP,
    HEUN2,
    RALSTON,
    HEUN3,
    RALSTON3,
    RK3,
    SSPRK3,
    HW,
    FEHLBERG2,
    FEHLBERG2_ADAPTIVE,
    RALSTON4,
    RULE38,
    BOSHA,
    BOSHA_ADAPTIVE,
    NYSTROEM,
    CASHKARP,
    CASHKARP_ADAPTIVE,
    FEHLBERG5,
    FEHLBERG5_ADAPTIVE,
    /* staged path only (more stages than the fused kernel keeps in registers) */
    NSRK78,
    NSRK78_ADAPTIVE,
    DP85,
    DP85_ADAPTIVE,
};

/* RungeKutta4::step (runge_kutta.cu:74-94) */
struct TabRK4 {
    static constexpr int S = 4, NB = 4;
    static constexpr bool ERR = false;
    static constexpr double a[4][4] = {{0, 0, 0, 0}, {0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1.0, 0}};
    static constexpr double c[4] = {0.0, 0.5, 0.5, 1.0};
    static constexpr double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    static constexpr double e[4] = {0, 0, 0, 0};
};

/* Dormand-Prince 5(4) (dormand_prince.cu:44-80); with ERR the 7th stage is evaluated
   on the 5th-order solution for the embedded error. */
struct TabDP45Base {
    static constexpr double a[7][7] = {
        {0, 0, 0, 0, 0, 0, 0},
        {1.0 / 5.0, 0, 0, 0, 0, 0, 0},
        {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0, 0},
        {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
        {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
        {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
        {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0}};
    static constexpr double c[7] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
    static constexpr double b[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0};
    static constexpr double e[7] = {35.0 / 384.0 - 5179.0 / 57600.0, 0.0, 500.0 / 1113.0 - 7571.0 / 16695.0, 125.0 / 192.0 - 393.0 / 640.0,
                                    -2187.0 / 6784.0 + 92097.0 / 339200.0, 11.0 / 84.0 - 187.0 / 2100.0, -1.0 / 40.0};
};
struct TabDP45Fixed : TabDP45Base {
    static constexpr int S = 6, NB = 6;
    static constexpr bool ERR = false;
};
struct TabDP45Adaptive : TabDP45Base {
    static constexpr int S = 7, NB = 6;
    static constexpr bool ERR = true;
};

/* Newton::step, explicit Euler (newton.cu:24-31) */
struct TabNewton {
    static constexpr int S = 1, NB = 1;
    static constexpr bool ERR = false;
    static constexpr double a[1][1] = {{0}};
    static constexpr double c[1] = {0.0};
    static constexpr double b[1] = {1.0};
    static constexpr double e[1] = {0};
};

/* MidPoint::step (explicit_mp.cu:28-39) */
struct TabMP {
    static constexpr int S = 2, NB = 2;
    static constexpr bool ERR = false;
    static constexpr double a[2][2] = {{0, 0}, {0.5, 0}};
    static constexpr double c[2] = {0.0, 0.5};
    static constexpr double b[2] = {0.0, 1.0};
    static constexpr double e[2] = {0, 0};
};

/* Heun2::step (heun.cu:31-42) */
struct TabHeun2 {
    static constexpr int S = 2, NB = 2;
    static constexpr bool ERR = false;
    static constexpr double a[2][2] = {{0, 0}, {1.0, 0}};
    static constexpr double c[2] = {0.0, 1.0};
    static constexpr double b[2] = {0.5, 0.5};
    static constexpr double e[2] = {0, 0};
};

/* Ralston::step (ralston.cu:31-42) */
struct TabRalston {
    static constexpr int S = 2, NB = 2;
    static constexpr bool ERR = false;
    static constexpr double a[2][2] = {{0, 0}, {2.0 / 3.0, 0}};
    static constexpr double c[2] = {0.0, 2.0 / 3.0};
    static constexpr double b[2] = {1.0 / 4.0, 3.0 / 4.0};
    static constexpr double e[2] = {0, 0};
};

/* Heun3::step (heun.cu:66-82) */
struct TabHeun3 {
    static constexpr int S = 3, NB = 3;
    static constexpr bool ERR = false;
    static constexpr double a[3][3] = {{0, 0, 0}, {1.0 / 3.0, 0, 0}, {0.0, 2.0 / 3.0, 0}};
    static constexpr double c[3] = {0.0, 1.0 / 3.0, 2.0 / 3.0};
    static constexpr double b[3] = {1.0 / 4.0, 0.0, 3.0 / 4.0};
    static constexpr double e[3] = {0, 0, 0};
};

/* Ralston3::step (ralston.cu:66-81) */
struct TabRalston3 {
    static constexpr int S = 3, NB = 3;
    static constexpr bool ERR = false;
    static constexpr double a[3][3] = {{0, 0, 0}, {1.0 / 2.0, 0, 0}, {0.0, 3.0 / 4.0, 0}};
    static constexpr double c[3] = {0.0, 1.0 / 2.0, 3.0 / 4.0};
    static constexpr double b[3] = {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0};
    static constexpr double e[3] = {0, 0, 0};
};

/* RungeKutta3::step (runge_kutta.cu:33-48), b as its tableau string (runge_kutta.cu:23-30) */
struct TabRK3 {
    static constexpr int S = 3, NB = 3;
    static constexpr bool ERR = false;
    static constexpr double a[3][3] = {{0, 0, 0}, {0.5, 0, 0}, {-1.0, 2.0, 0}};
    static constexpr double c[3] = {0.0, 0.5, 1.0};
    static constexpr double b[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
    static constexpr double e[3] = {0, 0, 0};
};

/* SSPRK3::step (SSPRK3.cu:27-42) */
struct TabSSPRK3 {
    static constexpr int S = 3, NB = 3;
    static constexpr bool ERR = false;
    static constexpr double a[3][3] = {{0, 0, 0}, {1.0, 0, 0}, {1.0 / 4.0, 1.0 / 4.0, 0}};
    static constexpr double c[3] = {0.0, 1.0, 1.0};
    static constexpr double b[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
    static constexpr double e[3] = {0, 0, 0};
};

/* HouwenWray::step (houwen_wray.cu:27-42) */
struct TabHW {
    static constexpr int S = 3, NB = 3;
    static constexpr bool ERR = false;
    static constexpr double a[3][3] = {{0, 0, 0}, {8.0 / 15.0, 0, 0}, {1.0 / 4.0, 5.0 / 12.0, 0}};
    static constexpr double c[3] = {0.0, 8.0 / 15.0, 2.0 / 3.0};
    static constexpr double b[3] = {1.0 / 4.0, 0.0, 3.0 / 4.0};
    static constexpr double e[3] = {0, 0, 0};
};

/* Fehlberg2::step (fehlberg.cu:33-70): fixed output, or output + error */
struct TabFehlberg2Base {
    static constexpr double a[3][3] = {{0, 0, 0}, {1.0 / 2.0, 0, 0}, {1.0 / 256.0, 255.0 / 256.0, 0}};
    static constexpr double c[3] = {0.0, 1.0 / 2.0, 1.0};
    static constexpr double b[3] = {1.0 / 512.0, 255.0 / 256.0, 1.0 / 512.0};
    static constexpr double e[3] = {1.0 / 512.0 - 1.0 / 256.0, 0.0, 1.0 / 512.0};
};
struct TabFehlberg2 : TabFehlberg2Base {
    static constexpr int S = 3, NB = 3;
    static constexpr bool ERR = false;
};
struct TabFehlberg2Adaptive : TabFehlberg2Base {
    static constexpr int S = 3, NB = 3;
    static constexpr bool ERR = true;
};

/* Ralston4::step (ralston.cu:106-125), decimal coefficients as given there */
struct TabRalston4 {
    static constexpr int S = 4, NB = 4;
    static constexpr bool ERR = false;
    static constexpr double a[4][4] = {{0, 0, 0, 0},
                                       {0.4, 0, 0, 0},
                                       {0.29697760924, 0.15875964497, 0, 0},
                                       {0.21810038822, -3.05096514869, 3.83286476047, 0}};
    static constexpr double c[4] = {0.0, 0.4, 0.51823725421, 0.51823725421};
    static constexpr double b[4] = {0.17476028226, -0.55148066287, 1.2055355994, 0.17118478122};
    static constexpr double e[4] = {0, 0, 0, 0};
};

/* Rule38::step (rule38.cu:28-48) */
struct TabRule38 {
    static constexpr int S = 4, NB = 4;
    static constexpr bool ERR = false;
    static constexpr double a[4][4] = {{0, 0, 0, 0}, {1.0 / 3.0, 0, 0, 0}, {-1.0 / 3.0, 1.0, 0, 0}, {1.0, -1.0, 1.0, 0}};
    static constexpr double c[4] = {0.0, 1.0 / 3.0, 2.0 / 3.0, 1.0};
    static constexpr double b[4] = {1.0 / 8.0, 3.0 / 8.0, 3.0 / 8.0, 1.0 / 8.0};
    static constexpr double e[4] = {0, 0, 0, 0};
};

/* BogackiShampine::step (bogacki.cu:28-75); the adaptive form evaluates a fourth
   stage on the 3rd-order output for the error (FSAL). */
struct TabBoShaBase {
    static constexpr double a[4][4] = {
        {0, 0, 0, 0}, {1.0 / 2.0, 0, 0, 0}, {0.0, 3.0 / 9.0, 0, 0}, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0, 0}};
    static constexpr double c[4] = {0.0, 1.0 / 2.0, 3.0 / 4.0, 1.0};
    static constexpr double b[4] = {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0, 0.0};
    static constexpr double e[4] = {2.0 / 9.0 - 7.0 / 24.0, 1.0 / 3.0 - 1.0 / 4.0, 4.0 / 9.0 - 1.0 / 3.0, 0.0 - 1.0 / 8.0};
};
struct TabBoSha : TabBoShaBase {
    static constexpr int S = 3, NB = 3;
    static constexpr bool ERR = false;
};
struct TabBoShaAdaptive : TabBoShaBase {
    static constexpr int S = 4, NB = 3;
    static constexpr bool ERR = true;
};

/* Nystroem::step (nystroem.cu:30-57) */
struct TabNystroem {
    static constexpr int S = 6, NB = 6;
    static constexpr bool ERR = false;
    static constexpr double a[6][6] = {{0, 0, 0, 0, 0, 0},
                                       {1.0 / 3.0, 0, 0, 0, 0, 0},
                                       {4.0 / 25.0, 6.0 / 25.0, 0, 0, 0, 0},
                                       {1.0 / 4.0, -3.0, 15.0 / 4.0, 0, 0, 0},
                                       {2.0 / 27.0, 10.0 / 9.0, -50.0 / 81.0, 8.0 / 81.0, 0, 0},
                                       {2.0 / 25.0, 12.0 / 25.0, 2.0 / 15.0, 8.0 / 75.0, 0, 0}};
    static constexpr double c[6] = {0.0, 1.0 / 3.0, 2.0 / 5.0, 1.0, 2.0 / 3.0, 4.0 / 5.0};
    static constexpr double b[6] = {23.0 / 192.0, 0.0, 125.0 / 192.0, 0.0, -27.0 / 64.0, 125.0 / 192.0};
    static constexpr double e[6] = {0, 0, 0, 0, 0, 0};
};

/* CashKarp::step (cash_karp.cu:31-83) */
struct TabCashKarpBase {
    static constexpr double a[6][6] = {{0, 0, 0, 0, 0, 0},
                                       {1.0 / 5.0, 0, 0, 0, 0, 0},
                                       {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
                                       {3.0 / 1.0, -9.0 / 10.0, 6.0 / 5.0, 0, 0, 0},
                                       {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0, 0},
                                       {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0, 0}};
    static constexpr double c[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
    static constexpr double b[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
    static constexpr double e[6] = {37.0 / 378.0 - 2825.0 / 27648.0, 0.0, 250.0 / 621.0 - 18575.0 / 48384.0,
                                    125.0 / 594.0 - 13525.0 / 55296.0, 0.0 - 2187.0 / 6784.0, 512.0 / 1771.0 - 1.0 / 4.0};
};
struct TabCashKarp : TabCashKarpBase {
    static constexpr int S = 6, NB = 6;
    static constexpr bool ERR = false;
};
struct TabCashKarpAdaptive : TabCashKarpBase {
    static constexpr int S = 6, NB = 6;
    static constexpr bool ERR = true;
};

/* Fehlberg5::step (fehlberg.cu:100-152) */
struct TabFehlberg5Base {
    static constexpr double a[6][6] = {{0, 0, 0, 0, 0, 0},
                                       {1.0 / 4.0, 0, 0, 0, 0, 0},
                                       {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0, 0},
                                       {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0, 0},
                                       {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0, 0},
                                       {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0, 0}};
    static constexpr double c[6] = {0.0, 1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
    static constexpr double b[6] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
    static constexpr double e[6] = {16.0 / 135.0 - 25.0 / 216.0, 0.0, 6656.0 / 12825.0 - 1408.0 / 2565.0,
                                    28561.0 / 56430.0 - 2197.0 / 4104.0, -9.0 / 50.0 - 1.0 / 5.0, 2.0 / 55.0};
};
struct TabFehlberg5 : TabFehlberg5Base {
    static constexpr int S = 6, NB = 6;
    static constexpr bool ERR = false;
};
struct TabFehlberg5Adaptive : TabFehlberg5Base {
    static constexpr int S = 6, NB = 6;
    static constexpr bool ERR = true;
};

/* NSRK78::step (NSRK78.cu:22-116): 13 stages, 8th-order output; the adaptive form sums the
   output into the buffer and the error with the weights its ERROR_K call passes (b minus a
   7th-order embedded solution), exponent 1/8.  Staged path only (phx_staged.hip). */
struct TabNSRK78Base {
    static constexpr double c[13] = {0.0, 0.001, 0.1111111111, 0.1666666667, 0.4166666667, 0.5, 0.8333333333, 0.1666666667, 0.6666666667, 0.76, 0.84, 1.0, 1.0};
    static constexpr double a[13][13] = {
        {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {1.0 / 1000.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {-491.0 / 81.0, 500.0 / 81.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {1.0 / 24.0, 0.0, 1.0 / 8.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {5.0 / 12.0, 0.0, -25.0 / 16.0, 25.0 / 16.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {1.0 / 20.0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 5.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {-43.0 / 180.0, 0.0, 0.0, 127.0 / 108.0, -329.0 / 135.0, 7.0 / 3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {27931.0 / 240300.0, 0.0, 0.0, -631.0 / 16020.0, 2459.0 / 6675.0, -3572.0 / 12015.0, 5.0 / 267.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {26114.0 / 12015.0, 0.0, 0.0, -7480.0 / 801.0, 67264.0 / 4005.0, -30640.0 / 2403.0, 1051.0 / 1335.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {33096587331.0 / 17382812500.0, 0.0, 0.0, -1828977848.0 / 173828125.0, 62801809904.0 / 4345703125.0, -9389764774.0 / 869140625.0, 6380757669.0 / 8691406250.0, 98417891.0 / 19531250.0, -1692691.0 / 39062500.0, 0.0, 0.0, 0.0, 0.0},
        {-1456295425347.0 / 2642187500000.0, 0.0, 0.0, 110740056.0 / 34765625.0, -21221682384.0 / 4345703125.0, 58859060169.0 / 13906250000.0, -177381525069.0 / 1529687500000.0, -28942485159.0 / 27812500000.0, -1272297.0 / 312500000.0, 5151.0 / 297616.0, 0.0, 0.0, 0.0},
        {844300798.0 / 137013275.0, 0.0, 0.0, -9997568.0 / 206035.0, 49636624.0 / 1030175.0, -3358834871.0 / 91067470.0, -40456983.0 / 1813108.0, 495817135.0 / 16647628.0, -149375.0 / 84266.0, 7470703125.0 / 1567431866.0, 1562500000.0 / 72342361.0, 0.0, 0.0},
        {-26225423.0 / 37371100.0, 0.0, 0.0, 807744.0 / 98345.0, -205584.0 / 37825.0, 5882202.0 / 1278485.0, -84543.0 / 432718.0, -223415.0 / 39338.0, -3625.0 / 6188.0, 292968750.0 / 374084711.0, 0.0, 0.0, 0.0}};
    static constexpr double b[13] = {4241.0 / 88200.0, 0.0, 0.0, 0.0, 0.0, 9946.0 / 23205.0, -10449.0 / 1925.0, 415449.0 / 1573075.0, -2025.0 / 5096.0, 48828125.0 / 44900856.0, 48828125.0 / 9843561.0, 463.0 / 12600.0, 0.0};
    static constexpr double e[13] = {4241.0 / 88200.0 - 3799.0 / 79800.0, 0.0, 0.0, 0.0, 0.0, 9946.0 / 23205.0 - 538.0 / 1365.0, -10449.0 / 1925.0 - 351.0 / 1925.0, 415449.0 / 1573075.0 - 4149.0 / 15575.0, -2025.0 / 5096.0 - (-45.0 / 392.0), 48828125.0 / 44900856.0 - 48828125.0 / 284372088.0, 48828125.0 / 9843561.0 - 0.0, 463.0 / 12600.0 - 0.0, 0.0 - 221.0 / 4200.0};
};
struct TabNSRK78 : TabNSRK78Base {
    static constexpr int S = 13, NB = 13;
    static constexpr bool ERR = false;
};
struct TabNSRK78Adaptive : TabNSRK78Base {
    static constexpr int S = 13, NB = 13;
    static constexpr bool ERR = true;
};

/* DormandPrince85::step (dormand_prince.cu:128-205): 12 stages, the coefficients of the
   reference's sequence.  The reference swaps its two branches (its fixed step only sums into
   the buffer, its adaptive step finalises in place and adapts on an error it never computed);
   here the fixed step applies b, and the adaptive step estimates its error with b - b5, b5 the
   5th-order weights its ERROR_K call passes, exponent 1/4 as passed to adaptTimeStep.  Staged
   path only. */
struct TabDP85Base {
    static constexpr double c[12] = {0.0, 0.05260015196, 0.07890022794, 0.1183503419, 0.2816496581, 0.3333333333, 0.25, 0.3076923077, 0.6512820513, 0.6, 0.8571428571, 1.0};
    static constexpr double a[12][12] = {
        {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {5.260015195876773e-2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {1.9725056984537899e-2, 5.9175170953613698e-2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {2.9587585476806849e-2, 0.0, 8.8762756430420548e-2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {2.4136513415926669e-1, 0.0, -8.8454947932828610e-1, 9.2483400326179200e-1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {3.7037037037037037e-2, 0.0, 0.0, 1.7082860872947387e-1, 1.2546768756682243e-1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {3.7109375e-2, 0.0, 0.0, 1.7025221101954404e-1, 6.0216538980455961e-2, -1.7578125e-2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {3.7092000118504793e-2, 0.0, 0.0, 1.7038392571223999e-1, 1.0726203044637328e-1, -1.5319437748624402e-2, 8.2737891638140229e-3, 0.0, 0.0, 0.0, 0.0, 0.0},
        {6.2411095871607572e-1, 0.0, 0.0, -3.3608926294469413e+0, -8.6821934684172601e-1, 2.7592099699446708e+1, 2.0154067550477893e+1, -4.3489884181069959e+1, 0.0, 0.0, 0.0, 0.0},
        {4.7766253643826437e-1, 0.0, 0.0, -2.4881146199716676e+0, -5.9029082683684300e-1, 2.1230051448181194e+1, 1.5279233632882424e+1, -3.3288210968984863e+1, -2.0331201708508626e-2, 0.0, 0.0, 0.0},
        {-9.3714243008598733e-1, 0.0, 0.0, 5.1863724288440637e+0, 1.0914373489967296e+0, -8.1497870107469261e+0, -1.8520065659996960e+1, 2.2739487099350504e+1, 2.4936055526796524e+0, -3.0467644718982195e+0, 0.0, 0.0},
        {2.2733101475165382e+0, 0.0, 0.0, -1.0534495466737250e+1, -2.0008720582248625e+0, -1.7958931863118799e+1, 2.7948884529419960e+1, -2.8589982771350237e+0, -8.8728569335306295e+0, 1.2360567175794303e+1, 6.4339274601576353e-1, 0.0}};
    static constexpr double b[12] = {5.4293734116568762e-2, 0.0, 0.0, 0.0, 0.0, 4.4503128927524089e+0, 1.8915178993145004e+0, -5.8012039600105848e+0, 3.1116436695781989e-1, -1.5216094966251608e-1, 2.0136540080403035e-1, 4.4710615727772591e-2};
    static constexpr double e[12] = {5.4293734116568762e-2 - (4.1173689122373888e-2), 0.0, 0.0, 0.0, 0.0, 4.4503128927524089e+0 - (5.6754693391286133e+0), 1.8915178993145004e+0 - (2.3872768489717506e+0), -5.8012039600105848e+0 - (-7.4655811424655713e+0), 3.1116436695781989e-1 - (6.6149321570779360e-1), -1.5216094966251608e-1 - (-4.8634006837553356e-1), 2.0136540080403035e-1 - (1.1944219431891464e-1), 4.4710615727772591e-2 - (6.7065923591658886e-2)};
};
struct TabDP85 : TabDP85Base {
    static constexpr int S = 12, NB = 12;
    static constexpr bool ERR = false;
};
struct TabDP85Adaptive : TabDP85Base {
    static constexpr int S = 12, NB = 12;
    static constexpr bool ERR = true;
};

/* Number of stages the fused kernel evaluates (= its halo in cells). */
inline int rk_tableau_stages(RkTableau t) {
    switch (t) {
        case RkTableau::NEWTON: return 1;
        case RkTableau::MP:
        case RkTableau::HEUN2:
        case RkTableau::RALSTON: return 2;
        case RkTableau::HEUN3:
        case RkTableau::RALSTON3:
        case RkTableau::RK3:
        case RkTableau::SSPRK3:
        case RkTableau::HW:
        case RkTableau::FEHLBERG2:
        case RkTableau::FEHLBERG2_ADAPTIVE:
        case RkTableau::BOSHA: return 3;
        case RkTableau::RK4:
        case RkTableau::RALSTON4:
        case RkTableau::RULE38:
        case RkTableau::BOSHA_ADAPTIVE: return 4;
        case RkTableau::DP45_FIXED:
        case RkTableau::NYSTROEM:
        case RkTableau::CASHKARP:
        case RkTableau::CASHKARP_ADAPTIVE:
        case RkTableau::FEHLBERG5:
        case RkTableau::FEHLBERG5_ADAPTIVE: return 6;
        case RkTableau::DP45_ADAPTIVE: return 7;
        case RkTableau::DP85:
        case RkTableau::DP85_ADAPTIVE: return 12;
        case RkTableau::NSRK78:
        case RkTableau::NSRK78_ADAPTIVE: return 13;
    }
    return 13;
}

/* The fused kernels keep every stage in registers and serve up to seven stages. */
inline bool rk_tableau_fused(RkTableau t) { return rk_tableau_stages(t) <= 7; }

/* Runtime form of a tableau, read by the staged path at launch time. */
constexpr int kStagedMaxStages = 13;
struct StagedTab {
    int S = 0, NB = 0;
    bool err = false;
    double c[kStagedMaxStages] = {}, b[kStagedMaxStages] = {}, e[kStagedMaxStages] = {};
    double a[kStagedMaxStages][kStagedMaxStages] = {};
};

template <class T>
inline StagedTab staged_from() {
    StagedTab t;
    t.S = T::S;
    t.NB = T::NB;
    t.err = T::ERR;
    for (int s = 0; s < T::S; s++) {
        t.c[s] = T::c[s];
        for (int j = 0; j < T::S; j++) t.a[s][j] = T::a[s][j];
        if (T::ERR) t.e[s] = T::e[s];
    }
    for (int j = 0; j < T::NB; j++) t.b[j] = T::b[j];
    return t;
}

inline StagedTab staged_tableau(RkTableau t) {
    switch (t) {
        case RkTableau::RK4: return staged_from<TabRK4>();
        case RkTableau::DP45_FIXED: return staged_from<TabDP45Fixed>();
        case RkTableau::DP45_ADAPTIVE: return staged_from<TabDP45Adaptive>();
        case RkTableau::NEWTON: return staged_from<TabNewton>();
        case RkTableau::MP: return staged_from<TabMP>();
        case RkTableau::HEUN2: return staged_from<TabHeun2>();
        case RkTableau::RALSTON: return staged_from<TabRalston>();
        case RkTableau::HEUN3: return staged_from<TabHeun3>();
        case RkTableau::RALSTON3: return staged_from<TabRalston3>();
        case RkTableau::RK3: return staged_from<TabRK3>();
        case RkTableau::SSPRK3: return staged_from<TabSSPRK3>();
        case RkTableau::HW: return staged_from<TabHW>();
        case RkTableau::FEHLBERG2: return staged_from<TabFehlberg2>();
        case RkTableau::FEHLBERG2_ADAPTIVE: return staged_from<TabFehlberg2Adaptive>();
        case RkTableau::RALSTON4: return staged_from<TabRalston4>();
        case RkTableau::RULE38: return staged_from<TabRule38>();
        case RkTableau::BOSHA: return staged_from<TabBoSha>();
        case RkTableau::BOSHA_ADAPTIVE: return staged_from<TabBoShaAdaptive>();
        case RkTableau::NYSTROEM: return staged_from<TabNystroem>();
        case RkTableau::CASHKARP: return staged_from<TabCashKarp>();
        case RkTableau::CASHKARP_ADAPTIVE: return staged_from<TabCashKarpAdaptive>();
        case RkTableau::FEHLBERG5: return staged_from<TabFehlberg5>();
        case RkTableau::FEHLBERG5_ADAPTIVE: return staged_from<TabFehlberg5Adaptive>();
        case RkTableau::NSRK78: return staged_from<TabNSRK78>();
        case RkTableau::NSRK78_ADAPTIVE: return staged_from<TabNSRK78Adaptive>();
        case RkTableau::DP85: return staged_from<TabDP85>();
        case RkTableau::DP85_ADAPTIVE: return staged_from<TabDP85Adaptive>();
    }
    return StagedTab{};
}

/* Iterator registry: the reference's REGISTER_SOLVER keys (solver_factory.hpp:47-60)
   served by the fused kernel.  adaptive == fixed for iterators registered without
   variable time stepping; power is the exponent of Solver::adaptTimeStep. */
struct IteratorInfo {
    const char* key;
    RkTableau fixed;
    RkTableau adaptive;
    bool can_adapt;
    double power;
};

inline const IteratorInfo* iterator_table(int* count) {
    static const IteratorInfo table[] = {
        {"Newton", RkTableau::NEWTON, RkTableau::NEWTON, false, 0.0},
        {"MP", RkTableau::MP, RkTableau::MP, false, 0.0},
        {"Heun2", RkTableau::HEUN2, RkTableau::HEUN2, false, 0.0},
        {"Ralston", RkTableau::RALSTON, RkTableau::RALSTON, false, 0.0},
        {"Heun3", RkTableau::HEUN3, RkTableau::HEUN3, false, 0.0},
        {"Ralston3", RkTableau::RALSTON3, RkTableau::RALSTON3, false, 0.0},
        {"RK3", RkTableau::RK3, RkTableau::RK3, false, 0.0},
        {"SSPRK3", RkTableau::SSPRK3, RkTableau::SSPRK3, false, 0.0},
        {"HW", RkTableau::HW, RkTableau::HW, false, 0.0},
        {"RK4", RkTableau::RK4, RkTableau::RK4, false, 0.0},
        {"Ralston4", RkTableau::RALSTON4, RkTableau::RALSTON4, false, 0.0},
        {"Rule38", RkTableau::RULE38, RkTableau::RULE38, false, 0.0},
        {"Nystroem", RkTableau::NYSTROEM, RkTableau::NYSTROEM, false, 0.0},
        {"Fehlberg2", RkTableau::FEHLBERG2, RkTableau::FEHLBERG2_ADAPTIVE, true, 1.0 / 4.0},
        {"BoSha", RkTableau::BOSHA, RkTableau::BOSHA_ADAPTIVE, true, 1.0 / 3.0},
        {"CashKarp", RkTableau::CASHKARP, RkTableau::CASHKARP_ADAPTIVE, true, 1.0 / 4.0},
        {"Fehlberg5", RkTableau::FEHLBERG5, RkTableau::FEHLBERG5_ADAPTIVE, true, 1.0 / 4.0},
        {"DP45", RkTableau::DP45_FIXED, RkTableau::DP45_ADAPTIVE, true, 1.0 / 4.0},
        {"DP85", RkTableau::DP85, RkTableau::DP85_ADAPTIVE, true, 1.0 / 4.0},
        {"NSRK78", RkTableau::NSRK78, RkTableau::NSRK78_ADAPTIVE, true, 1.0 / 8.0},
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

} // namespace phx
