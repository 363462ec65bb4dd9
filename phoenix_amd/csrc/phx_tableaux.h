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
    }
    return 7;
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
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

} // namespace phx
