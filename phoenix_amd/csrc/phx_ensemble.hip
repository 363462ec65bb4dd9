/*
 * phx_ensemble.hip -- batched ensembles of independent GP + reservoir grids of
 * different sizes (pump-power sweeps, noise seeds) stepped together.
 *
 * No reference counterpart: PHOENIX runs one system per process.  Here every
 * RK4 step of the whole ensemble is ONE fused launch over the flattened list of
 * all instances' wavefront tiles (rk4_batched_fused), and a run of steps is
 * captured once into a hipGraph and replayed, so small instances that would be
 * launch-bound on their own share the machine.  Arithmetic per cell is that of
 * the single-grid kernel, so an instance's result is bitwise the same as
 * stepping it alone.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/phx.h"
#include "phx_config.hpp"
#include "phx_device.h"
#include "phx_kernels.h"

namespace phx {
int set_error(int code, const std::string& msg); /* phx_solver.hip */
}

namespace {
int efail(int code, const std::string& m) { return phx::set_error(code, m); }
#define ENS_HIP(call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return efail(PHX_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)
} // namespace

struct phx_ensemble {
    int precision = PHX_FP32;
    size_t rsize = 4;
    int n = 0;
    std::vector<int> Nc, Nr;
    std::vector<double> dx, dy;
    int periodic_x = 1, periodic_y = 1;
    bool res = false, pump = false, pot = false;
    phx_physics ph{};
    void* pool = nullptr;
    /* per instance: offsets (bytes) of psi[2], n[2], pump, pot in the pool */
    std::vector<size_t> off_psi[2], off_n[2], off_pump, off_pot;
    int cur = 0;
    int ty = 12;
    bool ty_auto = true; /* tile height picked by timing at the first step */
    /* device tables for the two buffer parities and the tile list */
    void* entries[2] = {nullptr, nullptr};
    int* tile_inst = nullptr;
    int ntiles = 0;
    int tables_ty = -1;
    bool iso = false; /* all instances have dx == dy (set with the tables) */
    hipStream_t own_stream = nullptr, stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    double graph_dt = 0;
    int graph_ty = -1;
    int graph_parity = 0;

    char* base() const { return (char*)pool; }
};

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename real>
PhxPhys<real> phys_of(const phx_ensemble* e, int i) {
    phx::SystemConfig c;
    c.dx = e->dx[i];
    c.dy = e->dy[i];
    c.dV = c.dx * c.dy;
    c.one_over_dx2 = 1.0 / (c.dx * c.dx);
    c.one_over_dy2 = 1.0 / (c.dy * c.dy);
    c.m2_over_dx2_p_dy2 = -2.0 * (c.one_over_dx2 + c.one_over_dy2);
    c.m_eff_scaled = e->ph.m_eff_scaled;
    c.one_over_h_bar_s = e->ph.one_over_h_bar_s;
    c.gamma_c = e->ph.gamma_c;
    c.gamma_r = e->ph.gamma_r;
    c.g_c = e->ph.g_c;
    c.g_r = e->ph.g_r;
    c.R = e->ph.R;
    c.g_pm = e->ph.g_pm;
    c.delta_LT = e->ph.delta_LT;
    return phx::make_phys<real>(c);
}

/* (re)build the two entry tables and the tile list for tile height ty */
template <typename real>
int build_tables(phx_ensemble* e, int ty) {
    using v2 = typename phx::Vec2<real>::type;
    std::vector<phx::Rk4BatchEntry<real>> ent[2];
    std::vector<int> inst;
    e->iso = true;
    for (int i = 0; i < e->n; i++) {
        const PhxPhys<real> p = phys_of<real>(e, i);
        e->iso = e->iso && p.kin_yx == real(1);
        const int t0 = (int)inst.size();
        const int nt = phx::rk4_tiles(e->Nc[i], e->Nr[i], ty);
        for (int k = 0; k < nt; k++) inst.push_back(i);
        for (int par = 0; par < 2; par++) {
            phx::Rk4BatchEntry<real> en{};
            en.psi_in = (const v2*)(e->base() + e->off_psi[par][i]);
            en.psi_out = (v2*)(e->base() + e->off_psi[par ^ 1][i]);
            en.n_in = e->res ? (const real*)(e->base() + e->off_n[par][i]) : nullptr;
            en.n_out = e->res ? (real*)(e->base() + e->off_n[par ^ 1][i]) : nullptr;
            en.pump = e->pump ? (const real*)(e->base() + e->off_pump[i]) : nullptr;
            en.pot = e->pot ? (const real*)(e->base() + e->off_pot[i]) : nullptr;
            en.Nc = e->Nc[i];
            en.Nr = e->Nr[i];
            en.nstrips = phx::rk4_strips(e->Nc[i]);
            en.tile0 = t0;
            en.kin_c = p.kin_c;
            en.kin_x = p.kin_x;
            en.kin_y = p.kin_y;
            en.kin_yx = p.kin_yx;
            ent[par].push_back(en);
        }
    }
    /* launches still queued on the stream read the old tables: let them finish first */
    if (e->stream) ENS_HIP(hipStreamSynchronize(e->stream));
    for (int par = 0; par < 2; par++) {
        if (e->entries[par]) ENS_HIP(hipFree(e->entries[par]));
        ENS_HIP(hipMalloc(&e->entries[par], ent[par].size() * sizeof(ent[par][0])));
        ENS_HIP(hipMemcpy(e->entries[par], ent[par].data(), ent[par].size() * sizeof(ent[par][0]), hipMemcpyHostToDevice));
    }
    if (e->tile_inst) ENS_HIP(hipFree(e->tile_inst));
    ENS_HIP(hipMalloc((void**)&e->tile_inst, inst.size() * sizeof(int)));
    ENS_HIP(hipMemcpy(e->tile_inst, inst.data(), inst.size() * sizeof(int), hipMemcpyHostToDevice));
    e->ntiles = (int)inst.size();
    e->tables_ty = ty;
    return PHX_OK;
}

template <typename real>
int launch_step(phx_ensemble* e, double dt, int parity) {
    phx::Rk4BatchArgs<real> b{};
    b.entries = (const phx::Rk4BatchEntry<real>*)e->entries[parity];
    b.tile_inst = e->tile_inst;
    b.ntiles = e->ntiles;
    b.iso = e->iso;
    b.common.periodic_x = e->periodic_x;
    b.common.periodic_y = e->periodic_y;
    b.common.alt_rows = phx::rk4_alt_rows();
    b.common.dt = (real)dt;
    for (int k = 0; k < 4; k++) {
        b.common.pamp[k] = 1;
        b.common.vamp_h[k] = (real)e->ph.one_over_h_bar_s;
    }
    b.common.ph = phys_of<real>(e, 0);
    ENS_HIP(phx::launch_rk4_batched<real>(b, e->ty, e->res, e->pump, e->pot, e->stream));
    return PHX_OK;
}

/* Pick the tile height with the shortest step for this ensemble: every candidate
   launch reads the current buffers and writes the other ones, so the state
   does not advance.  Three round-robin passes, each candidate's fastest counting
   (a single pass in a fixed order times the first candidates while the clocks ramp). */
template <typename real>
int autotune(phx_ensemble* e, double dt) {
    hipEvent_t e0, e1;
    ENS_HIP(hipEventCreate(&e0));
    ENS_HIP(hipEventCreate(&e1));
    const int tys[] = {8, 10, 12, 16, 20};
    float ms_best[5] = {1e30f, 1e30f, 1e30f, 1e30f, 1e30f};
    int rc = PHX_OK;
    for (int pass = 0; pass < 3 && !rc; pass++) {
        for (int i = 0; i < 5 && !rc; i++) {
            rc = build_tables<real>(e, tys[i]);
            for (int w = 0; w < 2 && !rc; w++) rc = launch_step<real>(e, dt, e->cur);
            if (!rc && hipEventRecord(e0, e->stream) != hipSuccess) rc = efail(PHX_ERR_HIP, "hipEventRecord");
            for (int r = 0; r < 5 && !rc; r++) rc = launch_step<real>(e, dt, e->cur);
            if (!rc && hipEventRecord(e1, e->stream) != hipSuccess) rc = efail(PHX_ERR_HIP, "hipEventRecord");
            float ms = 0;
            if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = efail(PHX_ERR_HIP, "event timing");
            if (!rc) ms_best[i] = std::min(ms_best[i], ms);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    e->ty = tys[std::min_element(ms_best, ms_best + 5) - ms_best];
    e->ty_auto = false;
    return PHX_OK;
}

template <typename real>
int step_t(phx_ensemble* e, double dt, int64_t n_steps, int graph_steps) {
    if (e->ty_auto && n_steps > 0)
        if (int r = autotune<real>(e, dt)) return r;
    if (e->tables_ty != e->ty)
        if (int r = build_tables<real>(e, e->ty)) return r;
    int64_t done = 0;
    graph_steps &= ~1; /* even: the buffers return to the same parity */
    bool replay = graph_steps >= 2 && n_steps >= graph_steps;
    /* the captured launches hard-wire the buffer parity they start from */
    if (replay && (!e->graph || e->graph_steps != graph_steps || e->graph_dt != dt || e->graph_ty != e->ty ||
                   e->graph_parity != e->cur)) {
        if (e->graph) ENS_HIP(hipGraphExecDestroy(e->graph));
        e->graph = nullptr;
        if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            /* a stream that cannot be captured (the legacy default stream): launch by launch */
            (void)hipGetLastError();
            replay = false;
        } else {
            int rc = PHX_OK;
            for (int k = 0; k < graph_steps && !rc; k++) rc = launch_step<real>(e, dt, (e->cur + k) & 1);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(e->stream, &g);
            if (rc) return rc;
            ENS_HIP(ce);
            ENS_HIP(hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0));
            ENS_HIP(hipGraphDestroy(g));
            e->graph_steps = graph_steps;
            e->graph_dt = dt;
            e->graph_ty = e->ty;
            e->graph_parity = e->cur;
        }
    }
    if (replay) {
        while (n_steps - done >= graph_steps) {
            ENS_HIP(hipGraphLaunch(e->graph, e->stream));
            done += graph_steps;
        }
    }
    for (; done < n_steps; done++) {
        if (int r = launch_step<real>(e, dt, e->cur)) return r;
        e->cur ^= 1;
    }
    return PHX_OK;
}

int field_slot(phx_ensemble* e, int inst, int field, size_t* off, size_t* elems) {
    if (inst < 0 || inst >= e->n) return efail(PHX_ERR_INVALID, "instance out of range");
    const size_t N = (size_t)e->Nc[inst] * e->Nr[inst];
    switch (field) {
        case PHX_PSI_PLUS: *off = e->off_psi[e->cur][inst]; *elems = 2 * N; return PHX_OK;
        case PHX_N_PLUS:
            if (!e->res) break;
            *off = e->off_n[e->cur][inst];
            *elems = N;
            return PHX_OK;
        case PHX_PUMP_PLUS:
            if (!e->pump) break;
            *off = e->off_pump[inst];
            *elems = N;
            return PHX_OK;
        case PHX_POTENTIAL_PLUS:
            if (!e->pot) break;
            *off = e->off_pot[inst];
            *elems = N;
            return PHX_OK;
        default: break;
    }
    return efail(PHX_ERR_INVALID, "field not present in this ensemble");
}

} // namespace

extern "C" {

int phx_ensemble_create(phx_ensemble** out, int precision, const phx_physics* ph, int n, const int* N_c, const int* N_r, const double* dx,
                        const double* dy, int periodic_x, int periodic_y, int flags) {
    if (!out || !ph || n <= 0 || !N_c || !N_r) return efail(PHX_ERR_INVALID, "invalid argument");
    if (precision != PHX_FP32 && precision != PHX_FP64) return efail(PHX_ERR_INVALID, "precision");
    auto* e = new phx_ensemble();
    e->precision = precision;
    e->rsize = precision == PHX_FP32 ? 4 : 8;
    e->n = n;
    e->ph = *ph;
    e->periodic_x = periodic_x;
    e->periodic_y = periodic_y;
    e->res = flags & PHX_ENS_RESERVOIR;
    e->pump = (flags & PHX_ENS_PUMP) && e->res;
    e->pot = flags & PHX_ENS_POTENTIAL;
    size_t total = 0;
    for (int i = 0; i < n; i++) {
        if (N_c[i] <= 0 || N_r[i] <= 0) {
            delete e;
            return efail(PHX_ERR_INVALID, "instance " + std::to_string(i) + " has an empty grid");
        }
        if ((size_t)N_c[i] * (size_t)N_r[i] * 2 * e->rsize > phx::kMaxArrayBytes) {
            delete e;
            return efail(PHX_ERR_INVALID, "instance " + std::to_string(i) + " exceeds 2 GiB per field");
        }
        e->Nc.push_back(N_c[i]);
        e->Nr.push_back(N_r[i]);
        e->dx.push_back(dx ? dx[i] : ph->dx);
        e->dy.push_back(dy ? dy[i] : ph->dy);
        const size_t N = (size_t)N_c[i] * N_r[i];
        for (int b = 0; b < 2; b++) {
            e->off_psi[b].push_back(total);
            total = align256(total + 2 * N * e->rsize);
            e->off_n[b].push_back(total);
            if (e->res) total = align256(total + N * e->rsize);
        }
        e->off_pump.push_back(total);
        if (e->pump) total = align256(total + N * e->rsize);
        e->off_pot.push_back(total);
        if (e->pot) total = align256(total + N * e->rsize);
    }
    /* the pool is zeroed on the ensemble's own stream and waited for, so no later upload or step
       (on this or any other stream) can run before the fill has finished */
    if (hipMalloc(&e->pool, total) != hipSuccess || hipStreamCreate(&e->own_stream) != hipSuccess ||
        hipMemsetAsync(e->pool, 0, total, e->own_stream) != hipSuccess || hipStreamSynchronize(e->own_stream) != hipSuccess) {
        if (e->pool) (void)hipFree(e->pool);
        if (e->own_stream) (void)hipStreamDestroy(e->own_stream);
        delete e;
        return efail(PHX_ERR_HIP, "allocation of " + std::to_string(total) + " bytes failed");
    }
    e->stream = e->own_stream;
    *out = e;
    return PHX_OK;
}

int phx_ensemble_set_field(phx_ensemble* e, int inst, int field, const void* host, size_t bytes) {
    if (!e) return efail(PHX_ERR_INVALID, "null handle");
    size_t off, elems;
    if (int r = field_slot(e, inst, field, &off, &elems)) return r;
    if (bytes != elems * e->rsize) return efail(PHX_ERR_INVALID, "byte count mismatch");
    ENS_HIP(hipMemcpyAsync(e->base() + off, host, bytes, hipMemcpyHostToDevice, e->stream));
    ENS_HIP(hipStreamSynchronize(e->stream));
    return PHX_OK;
}

int phx_ensemble_get_field(phx_ensemble* e, int inst, int field, void* host, size_t bytes) {
    if (!e) return efail(PHX_ERR_INVALID, "null handle");
    size_t off, elems;
    if (int r = field_slot(e, inst, field, &off, &elems)) return r;
    if (bytes != elems * e->rsize) return efail(PHX_ERR_INVALID, "byte count mismatch");
    ENS_HIP(hipMemcpyAsync(host, e->base() + off, bytes, hipMemcpyDeviceToHost, e->stream));
    ENS_HIP(hipStreamSynchronize(e->stream));
    return PHX_OK;
}

int phx_ensemble_device_ptr(phx_ensemble* e, int inst, int field, void** dptr) {
    if (!e || !dptr) return efail(PHX_ERR_INVALID, "null argument");
    size_t off, elems;
    if (int r = field_slot(e, inst, field, &off, &elems)) return r;
    *dptr = e->base() + off;
    return PHX_OK;
}

int phx_ensemble_set_tile_rows(phx_ensemble* e, int ty) {
    if (!e) return efail(PHX_ERR_INVALID, "null handle");
    if (ty != 8 && ty != 10 && ty != 12 && ty != 16 && ty != 20) return efail(PHX_ERR_INVALID, "tile rows must be 8, 10, 12, 16 or 20");
    e->ty = ty;
    e->ty_auto = false;
    return PHX_OK;
}

int phx_ensemble_step(phx_ensemble* e, double dt, int64_t n_steps, int graph_steps) {
    if (!e || dt <= 0 || n_steps < 0) return efail(PHX_ERR_INVALID, "invalid argument");
    return e->precision == PHX_FP32 ? step_t<float>(e, dt, n_steps, graph_steps) : step_t<double>(e, dt, n_steps, graph_steps);
}

int phx_ensemble_set_stream(phx_ensemble* e, void* stream) {
    if (!e) return efail(PHX_ERR_INVALID, "null handle");
    e->stream = stream ? (hipStream_t)stream : e->own_stream;
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    return PHX_OK;
}

int phx_ensemble_sync(phx_ensemble* e) {
    if (!e) return efail(PHX_ERR_INVALID, "null handle");
    ENS_HIP(hipStreamSynchronize(e->stream));
    return PHX_OK;
}

void phx_ensemble_destroy(phx_ensemble* e) {
    if (!e) return;
    (void)hipStreamSynchronize(e->stream);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    for (void* p : e->entries)
        if (p) (void)hipFree(p);
    if (e->tile_inst) (void)hipFree(e->tile_inst);
    if (e->pool) (void)hipFree(e->pool);
    if (e->own_stream) (void)hipStreamDestroy(e->own_stream);
    delete e;
}

} /* extern "C" */
