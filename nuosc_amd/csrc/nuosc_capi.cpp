// nuosc_capi.cpp -- the extern "C" boundary (include/nuosc_hip.h): handle
// lifetime, device memory, the per-step launch sequence (optionally replayed
// from a hipGraph) and host<->device transfers.  Host C++; kernels live in
// nuosc_kernels.hip.
#include "../../include/nuosc_hip.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "nuosc_dev.h"
#include "vgrid.h"

using namespace nuosc;

namespace {
thread_local std::string g_create_error;
}

struct nuosc_sim {
    nuosc_config cfg{};
    // geometry (CartGrid.cpp:191-208)
    int nx = 0, nz = 0, nv = 0, nvp = 0, nch = 0;
    int nf = kNVar;                 // fields per site record: 8 (2 flavours) or 18 (3 flavours)
    long long srec = 0, xrow = 0;   // site-record and x-row strides of the state copies (elements)
    long long state_elems = 0;      // elements allocated for the four copies
    double dt = 0, ds_L = 0, ct = 0, st = 0;
    double phy_time = 0;
    double n_nue0[2] = {0, 0};
    std::vector<double> X, Z;
    VGrid vg;
    std::vector<double> G0h, G0bh;
    size_t esize = 8;  // bytes per stored value
    // device memory
    void* st_u = nullptr;
    void* st_w = nullptr;
    void* st_ya = nullptr;
    void* st_yb = nullptr;
    double* mom_part = nullptr;  // partial records of the last stage output
    double* mom_u = nullptr;     // totals of u, yA, yB (3 flavours: mom_u the stage input's, mom_a calRHS scratch)
    double* mom_a = nullptr;
    double* mom_b = nullptr;
    double* d_vgrid = nullptr;  // vx | vy | vz | vw
    double* d_g0 = nullptr;     // G0 | G0b
    double* d_part = nullptr;
    double* d_ana = nullptr;
    double* d_tmp = nullptr;    // FP64 staging (interior or ghosted field)
    double* d_time = nullptr;   // device copy of phy_time (open boundary data)
    Inflow inflow{};
    size_t tmp_bytes = 0;
    long long device_bytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph3 = nullptr;          // SSP-RK3 step
    hipGraphExec_t graph3_exec = nullptr;
    bool have_moments = false;
    int steps_in_last_call = 0;
    Geometry geo{};
    Physics ph{};
    StagePlan plan{};
    nuosc_batch* batch = nullptr;   // the batch this run is a member of, if any
    std::string err;
};

namespace {

int fail(nuosc_sim* s, int code, const std::string& msg) {
    if (s) s->err = msg;
    else g_create_error = msg;
    return code;
}

// A batch member changed what the batch's device tables hold (physics,
// decomposition): the tables are rebuilt before the next batch step.
void batch_member_changed(nuosc_batch* b);
// A member is being destroyed: it leaves its batch first.
void batch_detach(nuosc_batch* b, nuosc_sim* s);

#define HIPCK(sim, call)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((sim), NUOSC_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

bool is64(const nuosc_sim* s) { return s->cfg.precision == NUOSC_PREC_FP64; }

// Strides of the four state copies (separate allocations, site records of
// nf nvp elements); ok = false: the x rows one stage block addresses (2
// flavours: also the moment records; 3 flavours with nx > 1: the z march's
// block of rows) do not fit 32-bit buffer offsets.
struct CopyLayout {
    long long srec, xrow, elems;   // strides and elements allocated for all four copies
    bool ok;
};
CopyLayout copy_layout(int nx, int nz, int nzs, int nvp, int nch, int esize, int nf) {
    Geometry gq{};
    gq.nx = nx;
    gq.nz = nz;
    gq.nch = nch;
    gq.nvp = nvp;
    gq.nxs = nx + 2 * kXg + 1;
    gq.srec = (long long)nf * nvp;
    gq.xrow = (long long)nzs * gq.srec;
    if (nf == kNVar && !buffer_addressable(gq, esize)) return CopyLayout{0, 0, 0, false};
    if (nf == kNVar3 && !flavor3_addressable(gq, esize)) return CopyLayout{0, 0, 0, false};
    return CopyLayout{gq.srec, gq.xrow, 4 * copy_elems(gq), true};
}

void drop_graph(nuosc_sim* s) {
    // (teardown: errors of the destroy calls leave nothing to recover)
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    if (s->graph3_exec) (void)hipGraphExecDestroy(s->graph3_exec);
    if (s->graph3) (void)hipGraphDestroy(s->graph3);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    s->graph3_exec = nullptr;
    s->graph3 = nullptr;
}

int ensure_tmp(nuosc_sim* s, size_t bytes) {
    if (s->tmp_bytes >= bytes) return NUOSC_OK;
    if (s->d_tmp) {
        // the staging buffer may still be read by queued copies
        if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(s, NUOSC_EHIP, "stream sync failed");
        (void)hipFree(s->d_tmp);
        s->device_bytes -= (long long)s->tmp_bytes;
        s->d_tmp = nullptr;
        s->tmp_bytes = 0;
    }
    if (hipMalloc(&s->d_tmp, bytes) != hipSuccess) return fail(s, NUOSC_ENOMEM, "staging buffer allocation failed");
    s->tmp_bytes = bytes;
    s->device_bytes += (long long)bytes;
    return NUOSC_OK;
}

// 3-flavour vacuum Hamiltonian (nuosc_class.cpp:543-569), as written there:
// splittings times 1267*2, mixing angles degrees/180 used as radians,
// delta_cp = 0, minus a third of its trace on the diagonal.  Order: hee, hmm,
// htt, hemr, hemi, hmtr, hmti, hter, htei (the reference's names).
void vacuum_h3(double h[9]) {
    const double m21 = 7.39e-5 * 1267. * 2., m31 = 2.5229e-3 * 1267. * 2.;
    const double t12 = 33.82 / 180., t13 = 8.61 / 180., t23 = 48.3 / 180., cp = 0.;
    const double c12 = std::cos(t12), s12 = std::sin(t12), c13 = std::cos(t13), s13 = std::sin(t13);
    const double c23 = std::cos(t23), s23 = std::sin(t23);
    // elements of the mixing matrix used (real and imaginary parts)
    const double e2 = s12 * c13, e3r = s13 * std::cos(cp), e3i = -s13 * std::sin(cp);
    const double m2r = c12 * c23 - s12 * s13 * s23 * std::cos(cp), m2i = -s12 * s13 * s23 * std::sin(cp);
    const double u2r = -c12 * s23 - s12 * s13 * c23 * std::cos(cp), u2i = -s12 * s13 * c23 * std::sin(cp);
    const double m3 = c13 * s23, u3 = c13 * c23;
    const double dee = m21 * (e2 * e2) + m31 * (s13 * s13);
    const double dmm = m21 * (m2r * m2r + m2i * m2i) + m31 * (m3 * m3);
    const double dtt = m21 * (u2r * u2r + u2i * u2i) + m31 * (u3 * u3);
    h[3] = m21 * (e2 * m2r) + m31 * (e3r * m3);
    h[4] = m21 * (-e2 * m2i) + m31 * (e3i * m3);
    h[5] = m21 * (e2 * u2r) + m31 * (e3r * u3);
    h[6] = m21 * (-e2 * u2i) + m31 * (e3i * u3);
    h[7] = m21 * (m2r * u2r + m2i * u2i) + m31 * (m3 * u3);
    h[8] = m21 * (u2r * m2i - m2r * u2i);
    const double tr = (dee + dmm + dtt) / 3.0;
    h[0] = dee - tr;
    h[1] = dmm - tr;
    h[2] = dtt - tr;
}

// ghost rows of a stage input at stage time t + cdt: open z ends (if any),
// then the periodic x ghost rows (whole stored rows, z ghosts included)
template <typename T>
hipError_t stage_ghosts(nuosc_sim* s, T* y, double cdt) {
    if (s->geo.bc == NUOSC_BC_OPEN_Z) {
        const hipError_t e =
            launch_open_ghosts<T>(s->geo, y, s->d_g0, s->d_g0 + s->nv, s->inflow, s->d_time, cdt, s->stream);
        if (e != hipSuccess) return e;
    }
    // (one x row: the 3-flavour kernel's one-row form reads no x neighbours)
    return s->geo.nx == 1 ? hipSuccess : launch_xghosts<T>(s->geo, y, s->stream);
}

// 3 flavours: each stage is the ghost fill of its input, the stage kernel
// (the RK4 combination of enqueue_step; it reads the moment totals of its
// input from mom_u and writes the partial moments of its output), then the
// totals of the output into mom_u.  Events ev[2k], ev[2k+1] bracket the
// stage-(k+1) kernel alone.
template <typename T>
int enqueue_step3(nuosc_sim* s, hipEvent_t* ev) {
    const double dt = s->dt, c6 = 1.0 / 6.0 * dt, c2 = 0.5 * dt;
    T* u = static_cast<T*>(s->st_u);
    T* w = static_cast<T*>(s->st_w);
    T* ya = static_cast<T*>(s->st_ya);
    T* yb = static_cast<T*>(s->st_yb);
    const Geometry& g = s->geo;
    double* M = s->mom_u;
    double* P = s->mom_part;
    auto mark = [&](int k) { return ev ? hipEventRecord(ev[k], s->stream) : hipSuccess; };
    struct St { int stage; T* yin; T* uu; T* ww; T* yout; double cw, cy, ct; };
    const St st[4] = {{1, u, nullptr, nullptr, ya, 0.0, c2, 0.0},
                      {2, ya, u, w, yb, dt, c2, c2},
                      {3, yb, u, nullptr, ya, 0.0, dt, c2},
                      {4, ya, nullptr, w, u, c6, 0.0, dt}};
    for (int k = 0; k < 4; ++k) {
        HIPCK(s, stage_ghosts<T>(s, st[k].yin, st[k].ct));
        HIPCK(s, mark(2 * k));
        HIPCK(s, launch_stage3<T>(st[k].stage, g, s->ph, st[k].yin, st[k].uu, st[k].ww, st[k].yout, M, P,
                                  st[k].cw, st[k].cy, s->stream));
        HIPCK(s, mark(2 * k + 1));
        HIPCK(s, launch_moment_totals3(g, P, M, s->stream));
    }
    if (g.bc == NUOSC_BC_OPEN_Z) HIPCK(s, launch_advance_time(s->d_time, dt, s->stream));
    if (s->cfg.renorm) {
        HIPCK(s, launch_renorm<T>(g, u, s->d_g0, s->d_g0 + s->nv, s->stream));
        HIPCK(s, launch_moments3<T>(g, u, M, s->stream));
    }
    return NUOSC_OK;
}

// 3 flavours, SSP-RK3 (the stages of enqueue_step_srk3)
template <typename T>
int enqueue_step3_srk3(nuosc_sim* s) {
    const double dt = s->dt;
    T* u = static_cast<T*>(s->st_u);
    T* ya = static_cast<T*>(s->st_ya);
    T* yb = static_cast<T*>(s->st_yb);
    const Geometry& g = s->geo;
    double* M = s->mom_u;
    double* P = s->mom_part;
    HIPCK(s, stage_ghosts<T>(s, u, 0.0));
    HIPCK(s, launch_stage3<T>(5, g, s->ph, u, nullptr, nullptr, ya, M, P, 0.0, dt, s->stream));
    HIPCK(s, launch_moment_totals3(g, P, M, s->stream));
    HIPCK(s, stage_ghosts<T>(s, ya, dt));
    HIPCK(s, launch_stage3<T>(6, g, s->ph, ya, u, nullptr, yb, M, P, 0.0, dt, s->stream, 3.0 / 4.0, 1.0 / 4.0));
    HIPCK(s, launch_moment_totals3(g, P, M, s->stream));
    HIPCK(s, stage_ghosts<T>(s, yb, 0.5 * dt));
    HIPCK(s, launch_stage3<T>(7, g, s->ph, yb, u, nullptr, u, M, P, 0.0, dt, s->stream, 1.0 / 3.0, 2.0 / 3.0));
    HIPCK(s, launch_moment_totals3(g, P, M, s->stream));
    if (g.bc == NUOSC_BC_OPEN_Z) HIPCK(s, launch_advance_time(s->d_time, dt, s->stream));
    if (s->cfg.renorm) {
        HIPCK(s, launch_renorm<T>(g, u, s->d_g0, s->d_g0 + s->nv, s->stream));
        HIPCK(s, launch_moments3<T>(g, u, M, s->stream));
    }
    return NUOSC_OK;
}

// One RK4 step: the four stage launches (+ optional renormalisation).  With
// ev != nullptr events ev[2k], ev[2k+1] bracket the stage-(k+1) kernel alone
// (not its boundary fill or moment-total pass).
template <typename T>
int enqueue_step(nuosc_sim* s, hipEvent_t* ev = nullptr) {
    if (s->nf == kNVar3) return enqueue_step3<T>(s, ev);
    const double dt = s->dt;
    const double c6 = 1.0 / 6.0 * dt, c2 = 0.5 * dt;
    T* u = static_cast<T*>(s->st_u);
    T* w = static_cast<T*>(s->st_w);   // A = Y2 + u + dt k2 (see the stage kernel)
    T* ya = static_cast<T*>(s->st_ya);
    T* yb = static_cast<T*>(s->st_yb);
    const Geometry& g = s->geo;
    const bool comp = sizeof(T) == 4;
    double* P = s->mom_part;
    // ghost rows of each stage input: open z ends at the stage time t,
    // t + dt/2, t + dt/2, t + dt (t read from the device copy), then the
    // periodic x ghost rows (whole stored rows, z ghosts included)
    const bool open = g.bc == NUOSC_BC_OPEN_Z;
    auto ghosts = [&](T* y, double cdt) {
        if (open) {
            const hipError_t e =
                launch_open_ghosts<T>(g, y, s->d_g0, s->d_g0 + s->nv, s->inflow, s->d_time, cdt, s->stream);
            if (e != hipSuccess) return e;
        }
        // (the NOX stage kernel reads no x neighbours: no ghost rows needed)
        return g.nox ? hipSuccess : launch_xghosts<T>(g, y, s->stream);
    };
    auto mark = [&](int k) { return ev ? hipEventRecord(ev[k], s->stream) : hipSuccess; };
    HIPCK(s, ghosts(u, 0.0));
    HIPCK(s, mark(0));
    HIPCK(s, launch_stage<T>(1, g, s->ph, s->plan, u, nullptr, nullptr, ya, s->mom_u, P, 0.0, c2, s->stream));
    HIPCK(s, mark(1));
    HIPCK(s, launch_moment_totals(g, comp, P, s->mom_a, s->stream));
    HIPCK(s, ghosts(ya, c2));
    HIPCK(s, mark(2));
    HIPCK(s, launch_stage<T>(2, g, s->ph, s->plan, ya, u, w, yb, s->mom_a, P, dt, c2, s->stream));
    HIPCK(s, mark(3));
    HIPCK(s, launch_moment_totals(g, comp, P, s->mom_b, s->stream));
    HIPCK(s, ghosts(yb, c2));
    HIPCK(s, mark(4));
    HIPCK(s, launch_stage<T>(3, g, s->ph, s->plan, yb, u, nullptr, ya, s->mom_b, P, 0.0, dt, s->stream));
    HIPCK(s, mark(5));
    HIPCK(s, launch_moment_totals(g, comp, P, s->mom_a, s->stream));
    HIPCK(s, ghosts(ya, dt));
    HIPCK(s, mark(6));
    HIPCK(s, launch_stage<T>(4, g, s->ph, s->plan, ya, nullptr, w, u, s->mom_a, P, c6, 0.0, s->stream));
    HIPCK(s, mark(7));
    HIPCK(s, launch_moment_totals(g, comp, P, s->mom_u, s->stream));
    if (open) HIPCK(s, launch_advance_time(s->d_time, dt, s->stream));
    if (s->cfg.renorm) {
        HIPCK(s, launch_renorm<T>(g, u, s->d_g0, s->d_g0 + s->nv, s->stream));
        HIPCK(s, launch_moments<T>(g, u, P, s->stream));
        HIPCK(s, launch_moment_totals(g, comp, P, s->mom_u, s->stream));
    }
    return NUOSC_OK;
}

// One SSP-RK3 step (nuosc_class.cpp:429-452): u1 = u + dt L(u) -> yA,
// u2 = 3/4 u + 1/4 (u1 + dt L(u1)) -> yB, u' = 1/3 u + 2/3 (u2 + dt L(u2)) -> u,
// each stage input with its boundary fill (stage times t, t+dt, t+dt/2).
template <typename T>
int enqueue_step_srk3(nuosc_sim* s) {
    if (s->nf == kNVar3) return enqueue_step3_srk3<T>(s);
    const double dt = s->dt;
    T* u = static_cast<T*>(s->st_u);
    T* ya = static_cast<T*>(s->st_ya);
    T* yb = static_cast<T*>(s->st_yb);
    const Geometry& g = s->geo;
    const bool comp = sizeof(T) == 4;
    double* P = s->mom_part;
    const bool open = g.bc == NUOSC_BC_OPEN_Z;
    auto ghosts = [&](T* y, double cdt) {
        if (open) {
            const hipError_t e =
                launch_open_ghosts<T>(g, y, s->d_g0, s->d_g0 + s->nv, s->inflow, s->d_time, cdt, s->stream);
            if (e != hipSuccess) return e;
        }
        // (the NOX stage kernel reads no x neighbours: no ghost rows needed)
        return g.nox ? hipSuccess : launch_xghosts<T>(g, y, s->stream);
    };
    HIPCK(s, ghosts(u, 0.0));
    HIPCK(s, launch_stage<T>(5, g, s->ph, s->plan, u, nullptr, nullptr, ya, s->mom_u, P, 0.0, dt, s->stream));
    HIPCK(s, launch_moment_totals(g, comp, P, s->mom_a, s->stream));
    HIPCK(s, ghosts(ya, dt));
    HIPCK(s, launch_stage<T>(6, g, s->ph, s->plan, ya, u, nullptr, yb, s->mom_a, P, 0.0, dt, s->stream, 3.0 / 4.0,
                             1.0 / 4.0));
    HIPCK(s, launch_moment_totals(g, comp, P, s->mom_b, s->stream));
    HIPCK(s, ghosts(yb, 0.5 * dt));
    HIPCK(s, launch_stage<T>(7, g, s->ph, s->plan, yb, u, nullptr, u, s->mom_b, P, 0.0, dt, s->stream, 1.0 / 3.0,
                             2.0 / 3.0));
    HIPCK(s, launch_moment_totals(g, comp, P, s->mom_u, s->stream));
    if (open) HIPCK(s, launch_advance_time(s->d_time, dt, s->stream));
    if (s->cfg.renorm) {
        HIPCK(s, launch_renorm<T>(g, u, s->d_g0, s->d_g0 + s->nv, s->stream));
        HIPCK(s, launch_moments<T>(g, u, P, s->stream));
        HIPCK(s, launch_moment_totals(g, comp, P, s->mom_u, s->stream));
    }
    return NUOSC_OK;
}

// partial records and totals of the state u (3 flavours: the totals)
int refresh_moments(nuosc_sim* s) {
    if (s->nf == kNVar3) {
        if (is64(s)) HIPCK(s, launch_moments3<double>(s->geo, (double*)s->st_u, s->mom_u, s->stream));
        else HIPCK(s, launch_moments3<float>(s->geo, (float*)s->st_u, s->mom_u, s->stream));
        s->have_moments = true;
        return NUOSC_OK;
    }
    if (is64(s)) HIPCK(s, launch_moments<double>(s->geo, (double*)s->st_u, s->mom_part, s->stream));
    else HIPCK(s, launch_moments<float>(s->geo, (float*)s->st_u, s->mom_part, s->stream));
    HIPCK(s, launch_moment_totals(s->geo, !is64(s), s->mom_part, s->mom_u, s->stream));
    s->have_moments = true;
    return NUOSC_OK;
}

// Device analysis sums -> host (kNAna doubles).
int analysis_sums(nuosc_sim* s, double* h) {
    if (is64(s))
        HIPCK(s, launch_analysis<double>(s->geo, (double*)s->st_u, s->d_g0, s->d_g0 + s->nv, s->d_part, s->d_ana,
                                         s->stream));
    else
        HIPCK(s, launch_analysis<float>(s->geo, (float*)s->st_u, s->d_g0, s->d_g0 + s->nv, s->d_part, s->d_ana,
                                        s->stream));
    HIPCK(s, hipMemcpyAsync(h, s->d_ana, sizeof(double) * kNAna, hipMemcpyDeviceToHost, s->stream));
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

// ghost rows of state y at the current time: open z ends (if any), then the
// periodic x ghost rows
int fill_state_ghosts(nuosc_sim* s, void* y) {
    if (s->geo.bc == NUOSC_BC_OPEN_Z) {
        if (is64(s))
            HIPCK(s, launch_open_ghosts<double>(s->geo, (double*)y, s->d_g0, s->d_g0 + s->nv, s->inflow, s->d_time,
                                                0.0, s->stream));
        else
            HIPCK(s, launch_open_ghosts<float>(s->geo, (float*)y, s->d_g0, s->d_g0 + s->nv, s->inflow, s->d_time,
                                               0.0, s->stream));
    }
    if (is64(s)) HIPCK(s, launch_xghosts<double>(s->geo, (double*)y, s->stream));
    else HIPCK(s, launch_xghosts<float>(s->geo, (float*)y, s->stream));
    return NUOSC_OK;
}

int set_device_time(nuosc_sim* s) {
    if (!s->d_time) return NUOSC_OK;
    HIPCK(s, hipMemcpyAsync(s->d_time, &s->phy_time, sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

int check_field(nuosc_sim* s, int f) {
    if (f < 0 || f >= s->nf) return fail(s, NUOSC_EINVAL, "field index out of range");
    return NUOSC_OK;
}

// RK4 stages 1-3 with their boundary fills and moment totals: they write only
// the scratch copies A (w), yA, yB and the scratch moment records, never u.
template <typename T>
hipError_t enqueue_scratch_stages(nuosc_sim* s) {
    const double dt = s->dt, c2 = dt / 2;
    const Geometry& g = s->geo;
    T* u = static_cast<T*>(s->st_u);
    T* w = static_cast<T*>(s->st_w);
    T* ya = static_cast<T*>(s->st_ya);
    T* yb = static_cast<T*>(s->st_yb);
    const bool comp = sizeof(T) == 4, open = g.bc == NUOSC_BC_OPEN_Z;
    double* P = s->mom_part;
    hipError_t e = hipSuccess;
    auto ghosts = [&](T* y, double cdt) {
        if (e == hipSuccess && open)
            e = launch_open_ghosts<T>(g, y, s->d_g0, s->d_g0 + s->nv, s->inflow, s->d_time, cdt, s->stream);
        if (e == hipSuccess && !g.nox) e = launch_xghosts<T>(g, y, s->stream);
    };
    auto run = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    ghosts(u, 0.0);
    run(launch_stage<T>(1, g, s->ph, s->plan, u, nullptr, nullptr, ya, s->mom_u, P, 0.0, c2, s->stream));
    run(launch_moment_totals(g, comp, P, s->mom_a, s->stream));
    ghosts(ya, c2);
    run(launch_stage<T>(2, g, s->ph, s->plan, ya, u, w, yb, s->mom_a, P, dt, c2, s->stream));
    run(launch_moment_totals(g, comp, P, s->mom_b, s->stream));
    ghosts(yb, c2);
    run(launch_stage<T>(3, g, s->ph, s->plan, yb, u, nullptr, ya, s->mom_b, P, 0.0, dt, s->stream));
    return e;
}
}  // namespace

namespace nuosc {
SimView sim_view(const nuosc_sim* s) {
    return SimView{s->geo, s->st_u, s->d_g0, (int)s->esize, s->cfg.device, s->stream};
}
int sim_fail(nuosc_sim* s, int code, const char* msg) { return fail(s, code, msg); }
}  // namespace nuosc

extern "C" {

int nuosc_abi_version(void) { return NUOSC_ABI_VERSION; }

const char* nuosc_build_info(void) {
    return "libnuosc_hip: gfx950 RK4 stage kernels (fp64, fp32 + fp64 moments; periodic and open-z boundaries; 2 and 3 flavours), "
           "abi 5";
}

int nuosc_vgrid(int nn, int mode, double* vx, double* vy, double* vz, double* vw) {
    if (nn < 1 || nn > 64 || !vx || !vy || !vz || !vw) return -1;
    if (mode != NUOSC_VGRID_VORONOI && mode != NUOSC_VGRID_EQUAL) return -1;
    const VGrid g = (mode == NUOSC_VGRID_EQUAL) ? make_equal_weight_vgrid(nn) : make_icosahedral_vgrid(nn);
    const size_t b = sizeof(double) * g.size();
    std::memcpy(vx, g.vx.data(), b);
    std::memcpy(vy, g.vy.data(), b);
    std::memcpy(vz, g.vz.data(), b);
    std::memcpy(vw, g.vw.data(), b);
    return g.size();
}

const char* nuosc_last_error(const nuosc_sim* s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int nuosc_create(const nuosc_config* cfg, nuosc_sim** out) {
    if (!cfg || !out) return fail(nullptr, NUOSC_EINVAL, "null argument");
    *out = nullptr;
    const nuosc_config c = *cfg;
    if (c.nn < 1 || c.nn > 64) return fail(nullptr, NUOSC_EINVAL, "nn must be in [1, 64]");
    if (c.gx < 2 || c.gz < 2 || c.gx > 8 || c.gz > 8) return fail(nullptr, NUOSC_EINVAL, "ghost widths must be in [2, 8]");
    if (!(c.dx > 0) || !(c.dz > 0) || !(c.x1 > c.x0) || !(c.z1 > c.z0))
        return fail(nullptr, NUOSC_EINVAL, "bad box or spacing");
    if (c.precision != NUOSC_PREC_FP64 && c.precision != NUOSC_PREC_FP32)
        return fail(nullptr, NUOSC_EINVAL, "precision must be 64 or 32");
    if (c.bc != NUOSC_BC_PERIODIC && c.bc != NUOSC_BC_OPEN_Z)
        return fail(nullptr, NUOSC_EINVAL, "bc must be NUOSC_BC_PERIODIC or NUOSC_BC_OPEN_Z");
    if (c.bc == NUOSC_BC_OPEN_Z && c.gz > 3)
        return fail(nullptr, NUOSC_EINVAL, "the open z boundary takes gz = 2 or 3");
    if (c.ko_order != 0 && c.ko_order != 3 && c.ko_order != 5)
        return fail(nullptr, NUOSC_EINVAL, "ko_order must be 3 or 5");
    if (c.adv < NUOSC_ADV_CFD4 || c.adv > NUOSC_ADV_OFF) return fail(nullptr, NUOSC_EINVAL, "unknown advection scheme");
    if (c.bc == NUOSC_BC_OPEN_Z && c.gz < scheme_radius(c.ko_order == 5, c.adv))
        return fail(nullptr, NUOSC_EINVAL, "this discretisation has stencil radius 3: the open z boundary needs gz = 3");
    if (c.vgrid != NUOSC_VGRID_VORONOI && c.vgrid != NUOSC_VGRID_EQUAL)
        return fail(nullptr, NUOSC_EINVAL, "unknown velocity grid");
    if (c.nflavor != 0 && c.nflavor != 2 && c.nflavor != 3)
        return fail(nullptr, NUOSC_EINVAL, "nflavor must be 2 or 3");
    if (c.nflavor == 3 && (c.ko_order == 5 || c.adv != NUOSC_ADV_CFD4))
        return fail(nullptr, NUOSC_EINVAL, "3 flavours run with the default discretisation (KO 3, central advection)");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, NUOSC_ENODEV, "no HIP device");
    if (c.device < 0 || c.device >= ndev) return fail(nullptr, NUOSC_EINVAL, "device ordinal out of range");

    nuosc_sim* s = new nuosc_sim();
    s->cfg = c;
    s->nf = (c.nflavor == 3) ? kNVar3 : kNVar;
    // CartGrid.cpp:197-201
    s->nx = int((c.x1 - c.x0) / c.dx);
    s->nz = int((c.z1 - c.z0) / c.dz);
    if (s->nx < 1 || s->nz < 1) {
        delete s;
        return fail(nullptr, NUOSC_EINVAL, "empty grid");
    }
    if (c.bc == NUOSC_BC_OPEN_Z && s->nz < 3) {
        delete s;
        return fail(nullptr, NUOSC_EINVAL, "the open z boundary needs nz >= 3 (quadratic outflow extrapolation)");
    }
    const int zoff = (c.bc == NUOSC_BC_OPEN_Z) ? c.gz : 0;
    const int nzs = s->nz + 2 * zoff;
    s->X.resize(s->nx);
    s->Z.resize(s->nz);
    for (int i = 0; i < s->nx; ++i) s->X[i] = c.x0 + (i + 0.5) * c.dx;
    for (int i = 0; i < s->nz; ++i) s->Z[i] = c.z0 + (i + 0.5) * c.dz;
    s->vg = (c.vgrid == NUOSC_VGRID_EQUAL) ? make_equal_weight_vgrid(c.nn) : make_icosahedral_vgrid(c.nn);
    s->nv = s->vg.size();
    s->esize = is64(s) ? 8 : 4;
    const int align = (int)(128 / s->esize);
    s->nvp = (s->nv + align - 1) / align * align;
    const int lpr = dirs_per_chunk((int)s->esize);
    s->nch = (s->nv + lpr - 1) / lpr;
    {
        const CopyLayout L = copy_layout(s->nx, s->nz, nzs, s->nvp, s->nch, (int)s->esize, s->nf);
        if (!L.ok) {
            delete s;
            return fail(nullptr, NUOSC_EINVAL,
                        "grid too large for one GPU handle: a few stored x rows (nz * nv) or the moment records "
                        "(nx * nz * nv) exceed 4 GiB");
        }
        s->srec = L.srec;
        s->xrow = L.xrow;
        s->state_elems = L.elems;
    }
    // nuosc_class.cpp:461-464, nuosc_class.h:120-122
    s->ds_L = c.dx * c.dz / (c.z1 - c.z0) / (c.x1 - c.x0);
    s->dt = c.dz * c.cfl;
    const double theta = 37 * M_PI / 180.;
    s->ct = std::cos(2 * theta);
    s->st = std::sin(2 * theta);
    s->ph = Physics{c.mu, c.pmo, s->ct, s->st, {}};
    vacuum_h3(s->ph.h3);

    auto bail = [&](int code, const std::string& m) {
        std::string msg = m;
        nuosc_destroy(s);
        return fail(nullptr, code, msg);
    };
    if (hipSetDevice(c.device) != hipSuccess) return bail(NUOSC_EHIP, "hipSetDevice failed");
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(NUOSC_EHIP, "stream creation failed");
    if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
        return bail(NUOSC_EHIP, "event creation failed");

    const size_t state_bytes = (size_t)s->state_elems * s->esize;
    const bool three = s->nf == kNVar3;
    const size_t part_bytes = three ? (size_t)s->nx * s->nz * ((s->nv + 63) / 64) * kNMom3 * sizeof(double)
                                    : (size_t)(s->nx + 1) * s->nz * s->nch * kMomRec * sizeof(double);
    const size_t mom_bytes = (size_t)s->nx * s->nz * (three ? kNMom3 : kMomStride) * sizeof(double);
    void** states[4] = {&s->st_u, &s->st_w, &s->st_ya, &s->st_yb};
    {
        const size_t copy_bytes = state_bytes / 4;
        for (void** p : states) {
            if (hipMalloc(p, copy_bytes) != hipSuccess) return bail(NUOSC_ENOMEM, "state allocation failed");
            s->device_bytes += (long long)copy_bytes;
            if (hipMemsetAsync(*p, 0, copy_bytes, s->stream) != hipSuccess) return bail(NUOSC_EHIP, "memset failed");
        }
    }
    double** moms[4] = {&s->mom_part, &s->mom_u, &s->mom_a, &s->mom_b};
    for (double** p : moms) {
        if (three && p == &s->mom_b) continue;   // 3 flavours: totals of the stage input and a calRHS scratch
        const size_t b = (p == &s->mom_part) ? part_bytes : mom_bytes;
        if (hipMalloc((void**)p, b) != hipSuccess) return bail(NUOSC_ENOMEM, "moment allocation failed");
        s->device_bytes += (long long)b;
        if (hipMemsetAsync(*p, 0, b, s->stream) != hipSuccess) return bail(NUOSC_EHIP, "memset failed");
    }
    const size_t vbytes = 4 * (size_t)s->nv * sizeof(double);
    if (hipMalloc((void**)&s->d_vgrid, vbytes) != hipSuccess) return bail(NUOSC_ENOMEM, "allocation failed");
    if (hipMalloc((void**)&s->d_g0, 2 * s->nv * sizeof(double)) != hipSuccess)
        return bail(NUOSC_ENOMEM, "allocation failed");
    s->geo.nx = s->nx;
    s->geo.nz = s->nz;
    s->geo.nv = s->nv;
    s->geo.nvp = s->nvp;
    s->geo.nch = s->nch;
    s->geo.lpr = lpr;
    s->geo.rows = rows_per_wave((int)s->esize);
    s->geo.wpb = waves_per_block((int)s->esize);
    s->geo.nxs = s->nx + 2 * kXg + 1;
    s->geo.srec = s->srec;
    s->geo.xrow = s->xrow;
    s->geo.twelve_dx = 12 * c.dx;
    s->geo.twelve_dz = 12 * c.dz;
    s->geo.dx = c.dx;
    s->geo.dz = c.dz;
    s->geo.ko5 = c.ko_order == 5;
    s->geo.adv = c.adv;
    s->geo.ko_x = -c.ko / c.dx / 16.0;  // nuosc_class.cpp:120-121
    s->geo.ko_z = -c.ko / c.dz / 16.0;
    s->geo.vx = s->d_vgrid;
    s->geo.vy = s->d_vgrid + s->nv;
    s->geo.vz = s->d_vgrid + 2 * s->nv;
    s->geo.vw = s->d_vgrid + 3 * s->nv;
    s->geo.bc = c.bc;
    s->geo.nzs = nzs;
    s->geo.zoff = zoff;
    s->geo.nf = s->nf;
    // one x row with the default discretisation: the stage kernel's NOX form
    s->geo.nox = (s->nx == 1 && s->nf == kNVar && c.ko_order != 5 && c.adv == NUOSC_ADV_CFD4) ? 1 : 0;
    s->inflow = Inflow{0, 0.0, 1.0, c.zc, c.z0, c.dz};
    if (c.bc == NUOSC_BC_OPEN_Z) {
        if (hipMalloc((void**)&s->d_time, sizeof(double)) != hipSuccess) return bail(NUOSC_ENOMEM, "allocation failed");
        if (hipMemsetAsync(s->d_time, 0, sizeof(double), s->stream) != hipSuccess) return bail(NUOSC_EHIP, "memset failed");
    }
    const int nb = ana_blocks(s->geo);
    if (hipMalloc((void**)&s->d_part, (size_t)nb * kNAna * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&s->d_ana, kNAna * sizeof(double)) != hipSuccess)
        return bail(NUOSC_ENOMEM, "allocation failed");
    std::vector<double> hv(4 * (size_t)s->nv);
    for (int k = 0; k < s->nv; ++k) {
        hv[k] = s->vg.vx[k];
        hv[s->nv + k] = s->vg.vy[k];
        hv[2 * s->nv + k] = s->vg.vz[k];
        hv[3 * s->nv + k] = s->vg.vw[k];
    }
    if (hipMemcpyAsync(s->d_vgrid, hv.data(), vbytes, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return bail(NUOSC_EHIP, "copy failed");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c.device) != hipSuccess) return bail(NUOSC_EHIP, "device query failed");
    s->plan = plan_stage(s->geo, prop.multiProcessorCount);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(NUOSC_EHIP, "stream sync failed");
    s->G0h.assign(s->nv, 0.0);
    s->G0bh.assign(s->nv, 0.0);
    *out = s;
    return NUOSC_OK;
}

int nuosc_destroy(nuosc_sim* s) {
    if (!s) return NUOSC_OK;
    // a member leaves its batch before its buffers go (the batch's tables
    // and captured step point at them)
    if (s->batch) batch_detach(s->batch, s);
    // (teardown: a failed call here leaves nothing to recover, so the
    // remaining resources are released regardless)
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graph(s);
    void* bufs[] = {s->st_u, s->st_w, s->st_ya, s->st_yb, s->mom_part, s->mom_u, s->mom_a, s->mom_b,
                    s->d_vgrid, s->d_g0, s->d_part, s->d_ana, s->d_tmp, s->d_time};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return NUOSC_OK;
}

int nuosc_set_mu(nuosc_sim* s, double mu) {
    if (!s) return NUOSC_EINVAL;
    s->cfg.mu = mu;
    s->ph.mu = mu;
    drop_graph(s);
    if (s->batch) batch_member_changed(s->batch);
    return NUOSC_OK;
}

int nuosc_set_pmo(nuosc_sim* s, double pmo) {
    if (!s) return NUOSC_EINVAL;
    s->cfg.pmo = pmo;
    s->ph.pmo = pmo;
    drop_graph(s);
    if (s->batch) batch_member_changed(s->batch);
    return NUOSC_OK;
}

int nuosc_set_renorm(nuosc_sim* s, int renorm) {
    if (!s) return NUOSC_EINVAL;
    s->cfg.renorm = renorm ? 1 : 0;
    drop_graph(s);
    return NUOSC_OK;
}

int nuosc_get_dims(const nuosc_sim* s, int* nx, int* nz, int* nv, long long* lpts) {
    if (!s) return NUOSC_EINVAL;
    if (nx) *nx = s->nx;
    if (nz) *nz = s->nz;
    if (nv) *nv = s->nv;
    if (lpts) *lpts = (long long)(s->nx + 2 * s->cfg.gx) * (s->nz + 2 * s->cfg.gz) * s->nv;
    return NUOSC_OK;
}

int nuosc_get_scalars(const nuosc_sim* s, double* dt, double* ds_L, double* phy_time, double* n_nue0_2) {
    if (!s) return NUOSC_EINVAL;
    if (dt) *dt = s->dt;
    if (ds_L) *ds_L = s->ds_L;
    if (phy_time) *phy_time = s->phy_time;
    if (n_nue0_2) {
        n_nue0_2[0] = s->n_nue0[0];
        n_nue0_2[1] = s->n_nue0[1];
    }
    return NUOSC_OK;
}

int nuosc_set_phy_time(nuosc_sim* s, double t) {
    if (!s) return NUOSC_EINVAL;
    s->phy_time = t;
    return set_device_time(s);
}

int nuosc_get_vgrid(const nuosc_sim* s, double* vx, double* vy, double* vz, double* vw) {
    if (!s) return NUOSC_EINVAL;
    const size_t b = sizeof(double) * s->nv;
    if (vx) std::memcpy(vx, s->vg.vx.data(), b);
    if (vy) std::memcpy(vy, s->vg.vy.data(), b);
    if (vz) std::memcpy(vz, s->vg.vz.data(), b);
    if (vw) std::memcpy(vw, s->vg.vw.data(), b);
    return NUOSC_OK;
}

int nuosc_get_coords(const nuosc_sim* s, double* X, double* Z) {
    if (!s) return NUOSC_EINVAL;
    if (X) std::memcpy(X, s->X.data(), sizeof(double) * s->nx);
    if (Z) std::memcpy(Z, s->Z.data(), sizeof(double) * s->nz);
    return NUOSC_OK;
}

int nuosc_get_g0(const nuosc_sim* s, double* G0, double* G0b) {
    if (!s) return NUOSC_EINVAL;
    if (G0) std::memcpy(G0, s->G0h.data(), sizeof(double) * s->nv);
    if (G0b) std::memcpy(G0b, s->G0bh.data(), sizeof(double) * s->nv);
    return NUOSC_OK;
}

int nuosc_get_device_bytes(const nuosc_sim* s, long long* bytes) {
    if (!s || !bytes) return NUOSC_EINVAL;
    *bytes = s->device_bytes;
    return NUOSC_OK;
}

int nuosc_fill_init(nuosc_sim* s, int ipt, double alpha, double eps0, double sigma, double lnue, double lnueb,
                    double lnuex, double lnuebx, uint64_t seed) {
    if (!s) return NUOSC_EINVAL;
    if (!(ipt == 0 || ipt == 2 || ipt == 3 || ipt == 5))
        return fail(s, NUOSC_EINVAL, "ipt must be 0, 2, 3 or 5 (1 and 4 draw from rand() in the reference)");
    angular_distributions(s->vg, alpha, lnue, lnueb, lnuex, lnuebx, s->G0h, s->G0bh);
    // per-z perturbation (nuosc_init.cpp:4-6, 98-105)
    std::vector<double> tab(2 * (size_t)s->nz, 0.0);
    for (int j = 0; j < s->nz; ++j) {
        const double z = s->Z[j];
        double tmpr = 0.0;
        if (ipt == 0) tmpr = eps0 * std::exp(-(z - s->cfg.zc) * (z - s->cfg.zc) / (2.0 * sigma * sigma));
        else if (ipt == 2) tmpr = eps0 * (1.0 + std::cos(2 * M_PI * (z - 0.0) / (2.0 * sigma * sigma))) * 0.5;
        else if (ipt == 3) tmpr = eps0;
        tab[j] = tmpr;
        tab[s->nz + j] = std::sqrt(1.0 - tmpr * tmpr);
    }
    std::vector<double> g0(2 * (size_t)s->nv);
    std::memcpy(g0.data(), s->G0h.data(), sizeof(double) * s->nv);
    std::memcpy(g0.data() + s->nv, s->G0bh.data(), sizeof(double) * s->nv);
    HIPCK(s, hipMemcpyAsync(s->d_g0, g0.data(), sizeof(double) * 2 * s->nv, hipMemcpyHostToDevice, s->stream));
    int rc = ensure_tmp(s, sizeof(double) * 2 * (size_t)s->nz);
    if (rc) return rc;
    HIPCK(s, hipMemcpyAsync(s->d_tmp, tab.data(), sizeof(double) * 2 * s->nz, hipMemcpyHostToDevice, s->stream));
    if (is64(s))
        HIPCK(s, launch_init<double>(s->geo, ipt, s->d_g0, s->d_g0 + s->nv, s->d_tmp, s->d_tmp + s->nz, eps0, seed,
                                     (double*)s->st_u, s->stream));
    else
        HIPCK(s, launch_init<float>(s->geo, ipt, s->d_g0, s->d_g0 + s->nv, s->d_tmp, s->d_tmp + s->nz, eps0, seed,
                                    (float*)s->st_u, s->stream));
    // n_nue0 = ds_L * sum w ee (nuosc_init.cpp:116-130)
    double h[kNAna];
    rc = analysis_sums(s, h);
    if (rc) return rc;
    s->n_nue0[0] = h[ANA_SURV] * s->ds_L;
    s->n_nue0[1] = h[ANA_SURVB] * s->ds_L;
    s->phy_time = 0.0;
    // open z boundary: the inflow pulse continues the ipt-0 profile
    s->inflow.on = (ipt == 0) ? 1 : 0;
    s->inflow.eps0 = eps0;
    s->inflow.sigma = sigma;
    drop_graph(s);
    rc = set_device_time(s);
    if (rc) return rc;
    rc = refresh_moments(s);
    if (rc) return rc;
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

int nuosc_step_rk4(nuosc_sim* s, int nsteps) {
    if (!s) return NUOSC_EINVAL;
    if (nsteps < 0) return fail(s, NUOSC_EINVAL, "negative step count");
    if (!s->have_moments) {
        int rc = refresh_moments(s);
        if (rc) return rc;
    }
    HIPCK(s, hipEventRecord(s->ev0, s->stream));
    if (s->cfg.use_graph && nsteps > 0) {
        if (!s->graph_exec) {
            HIPCK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            int rc = is64(s) ? enqueue_step<double>(s) : enqueue_step<float>(s);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(s->stream, &gr);
            if (rc) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc;
            }
            if (e != hipSuccess) return fail(s, NUOSC_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
            s->graph = gr;
            HIPCK(s, hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0));
        }
        for (int n = 0; n < nsteps; ++n) {
            HIPCK(s, hipGraphLaunch(s->graph_exec, s->stream));
            s->phy_time += s->dt;
        }
    } else {
        for (int n = 0; n < nsteps; ++n) {
            int rc = is64(s) ? enqueue_step<double>(s) : enqueue_step<float>(s);
            if (rc) return rc;
            s->phy_time += s->dt;  // accumulated, as nuosc_class.cpp:423
        }
    }
    HIPCK(s, hipEventRecord(s->ev1, s->stream));
    s->steps_in_last_call = nsteps;
    return NUOSC_OK;
}

int nuosc_step_srk3(nuosc_sim* s, int nsteps) {
    if (!s) return NUOSC_EINVAL;
    if (nsteps < 0) return fail(s, NUOSC_EINVAL, "negative step count");
    if (!s->have_moments) {
        int rc = refresh_moments(s);
        if (rc) return rc;
    }
    HIPCK(s, hipEventRecord(s->ev0, s->stream));
    if (s->cfg.use_graph && nsteps > 0) {
        if (!s->graph3_exec) {
            HIPCK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            int rc = is64(s) ? enqueue_step_srk3<double>(s) : enqueue_step_srk3<float>(s);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(s->stream, &gr);
            if (rc) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc;
            }
            if (e != hipSuccess) return fail(s, NUOSC_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
            s->graph3 = gr;
            HIPCK(s, hipGraphInstantiate(&s->graph3_exec, s->graph3, nullptr, nullptr, 0));
        }
        for (int n = 0; n < nsteps; ++n) {
            HIPCK(s, hipGraphLaunch(s->graph3_exec, s->stream));
            s->phy_time += s->dt;
        }
    } else {
        for (int n = 0; n < nsteps; ++n) {
            int rc = is64(s) ? enqueue_step_srk3<double>(s) : enqueue_step_srk3<float>(s);
            if (rc) return rc;
            s->phy_time += s->dt;  // nuosc_class.cpp:448
        }
    }
    HIPCK(s, hipEventRecord(s->ev1, s->stream));
    s->steps_in_last_call = nsteps;
    return NUOSC_OK;
}

int nuosc_synchronize(nuosc_sim* s) {
    if (!s) return NUOSC_EINVAL;
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

int nuosc_last_step_ms(nuosc_sim* s, double* ms) {
    if (!s || !ms) return NUOSC_EINVAL;
    HIPCK(s, hipEventSynchronize(s->ev1));
    float f = 0;
    HIPCK(s, hipEventElapsedTime(&f, s->ev0, s->ev1));
    *ms = f;
    return NUOSC_OK;
}

int nuosc_profile_stages(nuosc_sim* s, int nsteps, double* stage_ms) {
    if (!s || !stage_ms || nsteps < 1) return NUOSC_EINVAL;
    if (!s->have_moments) {
        int rc = refresh_moments(s);
        if (rc) return rc;
    }
    hipEvent_t ev[8] = {};
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(s, NUOSC_EHIP, "event creation failed");
    double acc[4] = {0, 0, 0, 0};
    int rc = NUOSC_OK;
    for (int n = 0; n < nsteps && rc == NUOSC_OK; ++n) {
        rc = is64(s) ? enqueue_step<double>(s, ev) : enqueue_step<float>(s, ev);
        if (rc) break;
        s->phy_time += s->dt;
        if (hipEventSynchronize(ev[7]) != hipSuccess) {
            rc = fail(s, NUOSC_EHIP, "event sync failed");
            break;
        }
        for (int k = 0; k < 4 && rc == NUOSC_OK; ++k) {
            float f = 0;
            if (hipEventElapsedTime(&f, ev[2 * k], ev[2 * k + 1]) != hipSuccess)
                rc = fail(s, NUOSC_EHIP, "event timing failed");
            acc[k] += f;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 4; ++k) stage_ms[k] = acc[k] / nsteps;
    return rc;
}

int nuosc_get_plan(const nuosc_sim* s, int* zlen, int* waves) {
    if (!s) return NUOSC_EINVAL;
    if (zlen) *zlen = s->plan.zlen;
    if (waves) *waves = s->plan.ntask;
    return NUOSC_OK;
}

int nuosc_set_zlen(nuosc_sim* s, int zlen) {
    if (!s || zlen < 1) return NUOSC_EINVAL;
    HIPCK(s, hipStreamSynchronize(s->stream));
    s->plan = plan_stage_zlen(s->geo, zlen);
    drop_graph(s);
    if (s->batch) batch_member_changed(s->batch);
    return NUOSC_OK;
}

namespace {
std::mutex g_tune_mu;
// key: grid (nx, nz, nv), precision, z boundary, discretisation (KO order,
// advection scheme: they select the kernel instantiation) and the one-row form
std::map<std::tuple<int, int, int, int, int, int, int, int>, int> g_tune_cache;
}  // namespace


int nuosc_autotune(nuosc_sim* s, int trials, int* best_zlen, double* best_ms) {
    if (!s || trials < 1) return NUOSC_EINVAL;
    if (s->nf == kNVar3) {   // the 3-flavour kernels have no z-segment length to choose
        if (best_zlen) *best_zlen = 0;
        if (best_ms) *best_ms = -1;
        return NUOSC_OK;
    }
    const auto key = std::make_tuple(s->nx, s->nz, s->nv, s->cfg.precision, s->geo.bc, s->geo.ko5, s->geo.adv, s->geo.nox);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            int rc = nuosc_set_zlen(s, it->second);
            if (best_zlen) *best_zlen = it->second;
            if (best_ms) *best_ms = -1;
            return rc;
        }
    }
    if (!s->have_moments) {
        int rc = refresh_moments(s);
        if (rc) return rc;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(s, NUOSC_EHIP, "event creation failed");
    const StagePlan keep = s->plan;
    std::vector<int> cand;
    for (int zl = 8; zl <= 256; zl *= 2) {
        if (zl > s->nz && zl != 8) break;
        cand.push_back(zl);
    }
    std::vector<float> tmin(cand.size(), 1e30f);
    // candidates in turn, trials + 1 rounds (the first warms up), so a clock
    // ramp or drift during the search favours none of them
    int rc = NUOSC_OK;
    for (int t = 0; t < trials + 1 && rc == NUOSC_OK; ++t)
        for (size_t k = 0; k < cand.size() && rc == NUOSC_OK; ++k) {
            s->plan = plan_stage_zlen(s->geo, cand[k]);
            (void)hipEventRecord(e0, s->stream);
            const hipError_t le =
                is64(s) ? enqueue_scratch_stages<double>(s) : enqueue_scratch_stages<float>(s);
            (void)hipEventRecord(e1, s->stream);
            if (le != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(s, NUOSC_EHIP, "autotune launch failed");
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (t > 0) tmin[k] = std::min(tmin[k], ms);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        s->plan = keep;
        return rc;
    }
    size_t kb = 0;
    for (size_t k = 1; k < cand.size(); ++k)
        if (tmin[k] < tmin[kb]) kb = k;
    const int best = cand[kb];
    s->plan = plan_stage_zlen(s->geo, best);
    drop_graph(s);
    if (s->batch) batch_member_changed(s->batch);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune_cache[key] = best;
    }
    if (best_zlen) *best_zlen = best;
    if (best_ms) *best_ms = tmin[kb];
    return NUOSC_OK;
}

int nuosc_calrhs(nuosc_sim* s, const double* in, double* out) {
    if (!s || !in || !out) return NUOSC_EINVAL;
    const size_t fbytes = sizeof(double) * (size_t)s->nx * s->nz * s->nv;
    int rc = ensure_tmp(s, fbytes);
    if (rc) return rc;
    const size_t fe = (size_t)s->nx * s->nz * s->nv;
    for (int f = 0; f < s->nf; ++f) {
        HIPCK(s, hipMemcpyAsync(s->d_tmp, in + f * fe, fbytes, hipMemcpyHostToDevice, s->stream));
        if (is64(s)) HIPCK(s, launch_unpack<double>(s->geo, s->d_tmp, (double*)s->st_ya, f, s->stream));
        else HIPCK(s, launch_unpack<float>(s->geo, s->d_tmp, (float*)s->st_ya, f, s->stream));
    }
    rc = fill_state_ghosts(s, s->st_ya);
    if (rc) return rc;
    if (s->nf == kNVar3) {
        if (is64(s)) {
            HIPCK(s, launch_moments3<double>(s->geo, (double*)s->st_ya, s->mom_a, s->stream));
            HIPCK(s, launch_stage3<double>(0, s->geo, s->ph, (double*)s->st_ya, nullptr, nullptr, (double*)s->st_yb,
                                           s->mom_a, nullptr, 0.0, 0.0, s->stream));
        } else {
            HIPCK(s, launch_moments3<float>(s->geo, (float*)s->st_ya, s->mom_a, s->stream));
            HIPCK(s, launch_stage3<float>(0, s->geo, s->ph, (float*)s->st_ya, nullptr, nullptr, (float*)s->st_yb,
                                          s->mom_a, nullptr, 0.0, 0.0, s->stream));
        }
    } else if (is64(s)) {
        HIPCK(s, launch_moments<double>(s->geo, (double*)s->st_ya, s->mom_part, s->stream));
        HIPCK(s, launch_moment_totals(s->geo, false, s->mom_part, s->mom_a, s->stream));
        HIPCK(s, launch_stage<double>(0, s->geo, s->ph, s->plan, (double*)s->st_ya, nullptr, nullptr,
                                      (double*)s->st_yb, s->mom_a, nullptr, 0.0, 0.0, s->stream));
    } else {
        HIPCK(s, launch_moments<float>(s->geo, (float*)s->st_ya, s->mom_part, s->stream));
        HIPCK(s, launch_moment_totals(s->geo, true, s->mom_part, s->mom_a, s->stream));
        HIPCK(s, launch_stage<float>(0, s->geo, s->ph, s->plan, (float*)s->st_ya, nullptr, nullptr,
                                     (float*)s->st_yb, s->mom_a, nullptr, 0.0, 0.0, s->stream));
    }
    for (int f = 0; f < s->nf; ++f) {
        if (is64(s)) HIPCK(s, launch_pack<double>(s->geo, (double*)s->st_yb, f, s->d_tmp, s->stream));
        else HIPCK(s, launch_pack<float>(s->geo, (float*)s->st_yb, f, s->d_tmp, s->stream));
        HIPCK(s, hipMemcpyAsync(out + f * fe, s->d_tmp, fbytes, hipMemcpyDeviceToHost, s->stream));
    }
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

int nuosc_analysis(nuosc_sim* s, nuosc_analysis_t* o) {
    if (!s || !o) return NUOSC_EINVAL;
    double h[kNAna];
    int rc = analysis_sums(s, h);
    if (rc) return rc;
    // nuosc_ana.cpp:110-117
    const double nor = h[ANA_NOR], norb = h[ANA_NORB];
    o->phy_time = s->phy_time;
    o->maxdP = h[ANA_MAXDP];
    o->surv = h[ANA_SURV] / nor;
    o->survb = h[ANA_SURVB] / norb;
    o->avgP = h[ANA_AVGP] / nor;
    o->avgPb = h[ANA_AVGPB] / norb;
    o->aM0 = std::sqrt(h[ANA_AM01] * h[ANA_AM01] + h[ANA_AM02] * h[ANA_AM02] + h[ANA_AM03] * h[ANA_AM03]) * s->ds_L;
    o->ELNe = std::fabs(s->n_nue0[0] * (1.0 - o->surv) - s->n_nue0[1] * (1.0 - o->survb)) /
              (s->n_nue0[0] + s->n_nue0[1]);
    o->Lex = h[ANA_AM03] * s->ds_L;
    o->rho_ex2 = h[ANA_REX2] * s->ds_L;
    o->nonfinite = h[ANA_BAD];
    // nuosc_ana.cpp:131 asserts maxdP < 10; a NaN state also fails here
    const bool bad = !(o->maxdP < 10) || h[ANA_BAD] > 0 || !std::isfinite(o->surv) || !std::isfinite(o->aM0);
    o->status = bad ? NUOSC_EUNSTABLE : NUOSC_OK;
    if (bad) s->err = "analysis: state unstable (maxdP >= 10 or non-finite values)";
    return o->status;
}

int nuosc_renormalize(nuosc_sim* s) {
    if (!s) return NUOSC_EINVAL;
    if (is64(s)) HIPCK(s, launch_renorm<double>(s->geo, (double*)s->st_u, s->d_g0, s->d_g0 + s->nv, s->stream));
    else HIPCK(s, launch_renorm<float>(s->geo, (float*)s->st_u, s->d_g0, s->d_g0 + s->nv, s->stream));
    return refresh_moments(s);
}

int nuosc_download_field(nuosc_sim* s, int f, double* host) {
    if (!s || !host) return NUOSC_EINVAL;
    int rc = check_field(s, f);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * (size_t)s->nx * s->nz * s->nv;
    rc = ensure_tmp(s, bytes);
    if (rc) return rc;
    if (is64(s)) HIPCK(s, launch_pack<double>(s->geo, (double*)s->st_u, f, s->d_tmp, s->stream));
    else HIPCK(s, launch_pack<float>(s->geo, (float*)s->st_u, f, s->d_tmp, s->stream));
    HIPCK(s, hipMemcpyAsync(host, s->d_tmp, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

int nuosc_download_field_ghosted(nuosc_sim* s, int f, double* host) {
    if (!s || !host) return NUOSC_EINVAL;
    int rc = check_field(s, f);
    if (rc) return rc;
    const size_t bytes =
        sizeof(double) * (size_t)(s->nx + 2 * s->cfg.gx) * (s->nz + 2 * s->cfg.gz) * s->nv;
    rc = ensure_tmp(s, bytes);
    if (rc) return rc;
    rc = fill_state_ghosts(s, s->st_u);
    if (rc) return rc;
    if (is64(s))
        HIPCK(s, launch_ghosted<double>(s->geo, s->cfg.gx, s->cfg.gz, (double*)s->st_u, f, s->d_tmp,
                                        s->stream));
    else
        HIPCK(s, launch_ghosted<float>(s->geo, s->cfg.gx, s->cfg.gz, (float*)s->st_u, f, s->d_tmp,
                                       s->stream));
    HIPCK(s, hipMemcpyAsync(host, s->d_tmp, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

int nuosc_upload_field(nuosc_sim* s, int f, const double* host) {
    if (!s || !host) return NUOSC_EINVAL;
    int rc = check_field(s, f);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * (size_t)s->nx * s->nz * s->nv;
    rc = ensure_tmp(s, bytes);
    if (rc) return rc;
    HIPCK(s, hipMemcpyAsync(s->d_tmp, host, bytes, hipMemcpyHostToDevice, s->stream));
    if (is64(s)) HIPCK(s, launch_unpack<double>(s->geo, s->d_tmp, (double*)s->st_u, f, s->stream));
    else HIPCK(s, launch_unpack<float>(s->geo, s->d_tmp, (float*)s->st_u, f, s->stream));
    HIPCK(s, hipStreamSynchronize(s->stream));
    s->have_moments = false;
    return NUOSC_OK;
}

int nuosc_state_updated(nuosc_sim* s) {
    if (!s) return NUOSC_EINVAL;
    int rc = refresh_moments(s);
    if (rc) return rc;
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

int nuosc_p3_slice(nuosc_sim* s, const int* xidx, int sx, const int* vidx, int sv, double* dst) {
    if (!s || sx < 0 || sv < 0 || (sx && !xidx) || (sv && !vidx) || !dst) return NUOSC_EINVAL;
    for (int a = 0; a < sx; ++a)
        if (xidx[a] < 0 || xidx[a] >= s->nx) return fail(s, NUOSC_EINVAL, "x index out of range");
    for (int b = 0; b < sv; ++b)
        if (vidx[b] < 0 || vidx[b] >= s->nv) return fail(s, NUOSC_EINVAL, "v index out of range");
    const size_t n = (size_t)sx * s->nz * sv;
    if (n == 0) return NUOSC_OK;
    const size_t ibytes = sizeof(int) * (size_t)(sx + sv);
    const size_t ibytes_al = (ibytes + 255) / 256 * 256;
    int rc = ensure_tmp(s, ibytes_al + sizeof(double) * n);
    if (rc) return rc;
    int* di = reinterpret_cast<int*>(s->d_tmp);
    double* dd = reinterpret_cast<double*>(reinterpret_cast<char*>(s->d_tmp) + ibytes_al);
    std::vector<int> idx(xidx, xidx + sx);
    idx.insert(idx.end(), vidx, vidx + sv);
    HIPCK(s, hipMemcpyAsync(di, idx.data(), ibytes, hipMemcpyHostToDevice, s->stream));
    if (is64(s)) HIPCK(s, launch_p3_slice<double>(s->geo, (double*)s->st_u, s->d_g0, di, sx, di + sx, sv, dd, s->stream));
    else HIPCK(s, launch_p3_slice<float>(s->geo, (float*)s->st_u, s->d_g0, di, sx, di + sx, sv, dd, s->stream));
    HIPCK(s, hipMemcpyAsync(dst, dd, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
    HIPCK(s, hipStreamSynchronize(s->stream));
    return NUOSC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched runs
// ---------------------------------------------------------------------------
struct nuosc_batch {
    int device = 0;
    int precision = NUOSC_PREC_FP64;
    std::vector<nuosc_sim*> members;
    hipStream_t stream = nullptr;
    BatchTables tables;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // z-segment length shared by all members (0: each member's own plan).
    // One length gives every block of the batched launch the same amount of
    // work, so the contiguous block ranges of the 8 XCDs stay balanced.
    int zlen = 64;
    int launch = kBatchGrid;   // NUOSC_BATCH_GRID or NUOSC_BATCH_PERSISTENT
    std::string err;
};
static_assert(NUOSC_BATCH_GRID == kBatchGrid && NUOSC_BATCH_PERSISTENT == kBatchPersistent,
              "launch forms of the header and the kernels agree");

namespace {
int bfail(nuosc_batch* b, int code, const std::string& m) {
    b->err = m;
    return code;
}
void batch_drop_graph(nuosc_batch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}
void batch_member_changed(nuosc_batch* b) { b->dirty = true; }
void batch_detach(nuosc_batch* b, nuosc_sim* s) {
    (void)hipStreamSynchronize(b->stream);
    b->members.erase(std::remove(b->members.begin(), b->members.end(), s), b->members.end());
    s->batch = nullptr;
    b->dirty = true;
    batch_drop_graph(b);
}
}  // namespace

extern "C" {

namespace {
// (re)build the device tables of the batched launches from the members
int batch_tables(nuosc_batch* b) {
    std::vector<BatchMember> ms;
    for (nuosc_sim* m : b->members) {
        BatchMember bm;
        bm.g = m->geo;
        bm.g.nox = 0;   // the batched stage kernel has no NOX form
        bm.ph = m->ph;
        bm.pl = plan_stage_zlen(bm.g, b->zlen > 0 ? std::min(b->zlen, m->nz) : m->plan.zlen);
        bm.u = m->st_u;
        bm.w = m->st_w;
        bm.ya = m->st_ya;
        bm.yb = m->st_yb;
        bm.mpart = m->mom_part;
        bm.mu = m->mom_u;
        bm.ma = m->mom_a;
        bm.mb = m->mom_b;
        bm.dt = m->dt;
        ms.push_back(bm);
    }
    hipError_t e = batch_build(b->tables, ms, is64(b->members[0]) ? 8 : 4, b->launch, b->stream);
    if (e != hipSuccess) return bfail(b, NUOSC_EHIP, std::string("batch tables: ") + hipGetErrorString(e));
    return NUOSC_OK;
}

// members' own queued work (init, uploads, moments) completes first
int batch_sync_members(nuosc_batch* b) {
    for (nuosc_sim* m : b->members) {
        if (m->cfg.renorm) return bfail(b, NUOSC_EINVAL, "renormalising runs cannot be batched");
        if (!m->have_moments && refresh_moments(m) != NUOSC_OK) return bfail(b, NUOSC_EHIP, m->err);
        if (hipStreamSynchronize(m->stream) != hipSuccess) return bfail(b, NUOSC_EHIP, "member stream sync failed");
    }
    return NUOSC_OK;
}
}  // namespace

int nuosc_batch_create(int device, int precision, nuosc_batch** out) {
    if (!out) return NUOSC_EINVAL;
    *out = nullptr;
    if (precision != NUOSC_PREC_FP64 && precision != NUOSC_PREC_FP32)
        return fail(nullptr, NUOSC_EINVAL, "precision must be 64 or 32");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, NUOSC_ENODEV, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, NUOSC_EINVAL, "device ordinal out of range");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, NUOSC_EHIP, "hipSetDevice failed");
    nuosc_batch* b = new nuosc_batch();
    b->device = device;
    b->precision = precision;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(nullptr, NUOSC_EHIP, "stream creation failed");
    }
    *out = b;
    return NUOSC_OK;
}

int nuosc_batch_destroy(nuosc_batch* b) {
    if (!b) return NUOSC_OK;
    (void)hipStreamSynchronize(b->stream);
    for (nuosc_sim* m : b->members) m->batch = nullptr;
    batch_drop_graph(b);
    batch_free(b->tables);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return NUOSC_OK;
}

const char* nuosc_batch_last_error(const nuosc_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int nuosc_batch_add(nuosc_batch* b, nuosc_sim* s) {
    if (!b || !s) return NUOSC_EINVAL;
    if (s->cfg.device != b->device || s->cfg.precision != b->precision)
        return bfail(b, NUOSC_EINVAL, "member must use the batch's device and precision");
    if (s->cfg.renorm) return bfail(b, NUOSC_EINVAL, "renormalising runs cannot be batched");
    if (s->cfg.bc != NUOSC_BC_PERIODIC) return bfail(b, NUOSC_EINVAL, "only periodic runs can be batched");
    if (s->geo.ko5 || s->geo.adv != NUOSC_ADV_CFD4)
        return bfail(b, NUOSC_EINVAL, "only runs with the default discretisation can be batched");
    if (s->nf != kNVar) return bfail(b, NUOSC_EINVAL, "only 2-flavour runs can be batched");
    if (s->batch == b) return bfail(b, NUOSC_EINVAL, "already a member");
    if (s->batch) return bfail(b, NUOSC_EINVAL, "already a member of another batch");
    b->members.push_back(s);
    s->batch = b;
    b->dirty = true;
    return NUOSC_OK;
}

int nuosc_batch_remove(nuosc_batch* b, nuosc_sim* s) {
    if (!b || !s) return NUOSC_EINVAL;
    for (size_t k = 0; k < b->members.size(); ++k)
        if (b->members[k] == s) {
            batch_detach(b, s);
            return NUOSC_OK;
        }
    return bfail(b, NUOSC_EINVAL, "not a member");
}

int nuosc_batch_size(const nuosc_batch* b) { return b ? (int)b->members.size() : 0; }

int nuosc_batch_step(nuosc_batch* b, int nsteps) {
    if (!b || nsteps < 0) return NUOSC_EINVAL;
    if (b->members.empty() || nsteps == 0) return NUOSC_OK;
    int rc0 = batch_sync_members(b);
    if (rc0) return rc0;
    if (b->dirty) {
        batch_drop_graph(b);
        int rc = batch_tables(b);
        if (rc) return rc;
        if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return bfail(b, NUOSC_EHIP, "capture failed");
        hipError_t le = batch_step(b->tables, b->stream);
        hipGraph_t gr = nullptr;
        hipError_t ce = hipStreamEndCapture(b->stream, &gr);
        if (le != hipSuccess || ce != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            return bfail(b, NUOSC_EHIP, "batched step capture failed");
        }
        b->graph = gr;
        if (hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
            return bfail(b, NUOSC_EHIP, "graph instantiation failed");
        b->dirty = false;
    }
    for (int n = 0; n < nsteps; ++n) {
        if (hipGraphLaunch(b->graph_exec, b->stream) != hipSuccess) return bfail(b, NUOSC_EHIP, "graph launch failed");
        for (nuosc_sim* m : b->members) m->phy_time += m->dt;
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bfail(b, NUOSC_EHIP, "batch stream sync failed");
    return NUOSC_OK;
}

int nuosc_batch_set_zlen(nuosc_batch* b, int zlen) {
    if (!b || zlen < 0) return NUOSC_EINVAL;
    (void)hipStreamSynchronize(b->stream);
    b->zlen = zlen;
    b->dirty = true;
    return NUOSC_OK;
}

int nuosc_batch_set_launch(nuosc_batch* b, int launch) {
    if (!b || (launch != NUOSC_BATCH_GRID && launch != NUOSC_BATCH_PERSISTENT)) return NUOSC_EINVAL;
    (void)hipStreamSynchronize(b->stream);
    b->launch = launch;
    b->dirty = true;
    return NUOSC_OK;
}

int nuosc_batch_get_launch(const nuosc_batch* b) { return b ? b->launch : -1; }

int nuosc_batch_autotune(nuosc_batch* b, int trials, int* best_zlen, double* best_ms) {
    if (!b || trials < 1) return NUOSC_EINVAL;
    if (b->members.empty()) return bfail(b, NUOSC_EINVAL, "empty batch");
    int rc = batch_sync_members(b);
    if (rc) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return bfail(b, NUOSC_EHIP, "event creation failed");
    const int keep = b->zlen, keep_launch = b->launch;
    // (z-segment length, launch form) pairs
    std::vector<std::pair<int, int>> cand;
    for (int zl : {16, 32, 64, 128, 256})
        for (int lf : {NUOSC_BATCH_GRID, NUOSC_BATCH_PERSISTENT}) cand.push_back({zl, lf});
    std::vector<float> tmin(cand.size(), 1e30f);
    // candidates in turn (tables rebuilt per candidate), trials + 1 rounds,
    // the first a warm-up, so clock drift during the search favours none
    for (int t = 0; t < trials + 1 && rc == NUOSC_OK; ++t)
        for (size_t k = 0; k < cand.size() && rc == NUOSC_OK; ++k) {
            b->zlen = cand[k].first;
            b->launch = cand[k].second;
            rc = batch_tables(b);
            if (rc) break;
            (void)hipEventRecord(e0, b->stream);
            // stage 1 writes only the RK4 scratch copies (w, yA), never the state
            hipError_t le = batch_stage(b->tables, 1, b->stream);
            (void)hipEventRecord(e1, b->stream);
            if (le != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = bfail(b, NUOSC_EHIP, "autotune launch failed");
                break;
            }
            float ms = 0;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = bfail(b, NUOSC_EHIP, "autotune timing failed");
                break;
            }
            if (t > 0) tmin[k] = std::min(tmin[k], ms);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    size_t kb = 0;
    for (size_t k = 1; k < cand.size(); ++k)
        if (tmin[k] < tmin[kb]) kb = k;
    b->zlen = rc == NUOSC_OK ? cand[kb].first : keep;
    b->launch = rc == NUOSC_OK ? cand[kb].second : keep_launch;
    b->dirty = true;
    batch_drop_graph(b);
    if (rc) return rc;
    if (best_zlen) *best_zlen = cand[kb].first;
    if (best_ms) *best_ms = tmin[kb];
    return NUOSC_OK;
}

int nuosc_estimate_device_bytes(const nuosc_config* c, long long* bytes) {
    if (!c || !bytes || c->nn < 1 || c->nn > 64 || !(c->dx > 0) || !(c->dz > 0)) return NUOSC_EINVAL;
    const long long nx = (long long)((c->x1 - c->x0) / c->dx), nz = (long long)((c->z1 - c->z0) / c->dz);
    if (nx < 1 || nz < 1) return NUOSC_EINVAL;
    const long long nv = 10ll * c->nn * c->nn + 2;
    const int esize = c->precision == NUOSC_PREC_FP32 ? 4 : 8;
    const long long align = 128 / esize, nvp = (nv + align - 1) / align * align;
    const long long lpr = dirs_per_chunk(esize), nch = (nv + lpr - 1) / lpr;
    const long long nzs = nz + (c->bc == NUOSC_BC_OPEN_Z ? 2 * c->gz : 0);
    const int nf = c->nflavor == 3 ? kNVar3 : kNVar;
    const CopyLayout L = copy_layout((int)nx, (int)nz, (int)nzs, (int)nvp, (int)nch, esize, nf);
    if (!L.ok) return NUOSC_EINVAL;
    long long b = L.elems * esize;                                     // u, w, yA, yB
    if (nf == kNVar3) b += nx * nz * ((nv + 63) / 64 + 2) * kNMom3 * 8;   // moments: partials, 2 totals
    else b += (nx + 1) * nz * nch * kMomRec * 8 + 3ll * nx * nz * kMomStride * 8;
    b += 4 * nv * 8 + 2 * nv * 8 + 2048ll * kNAna * 16 + kNAna * 8;    // v-grid, G0, analysis
    b += 2 * nz * 8;                                                   // init tables
    *bytes = b;
    return NUOSC_OK;
}

}  // extern "C"
