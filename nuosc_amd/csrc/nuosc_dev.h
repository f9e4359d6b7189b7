// nuosc_dev.h -- internal interface between the C-ABI layer (nuosc_capi.cpp)
// and the HIP kernels (nuosc_kernels.hip).  Not part of the public ABI.
//
// Device data layout (state copies u, w, yA, yB):
//   site records of nf fields, each field a row of nvp directions (v
//   fastest; nvp = nv rounded up so that every field row starts on a 128-byte
//   boundary):
//     copy[xs][zs][f][nvp],  xs = i + kXg (stored x row), zs = j + zoff
//   so the fields of a site are one contiguous block (srec = nf nvp elements
//   between site records) and a stored x row (nzs records) is contiguous.
//   2 flavours: nf = 8 (ee, mm, emr, emi, bee, bmm, bemr, bemi); 3 flavours:
//   nf = 18, the same 8 first, then tt, mtr, mti, ter, tei, btt, bmtr, bmti,
//   bter, btei.
//   Each copy is its own allocation.  Stored x rows: kXg = 3 periodic
//   ghost rows at each end (copies of interior rows nx-3 .. nx-1 and
//   0 .. 2, refreshed before every stage from the interior) and one spare row at
//   xs = nx + 2 kXg, where stage-kernel lanes past the last x row store
//   instead of branching; nxs = nx + 2 kXg + 1.  The x-neighbours of every
//   wave are therefore one contiguous address range (a few stored rows, one
//   buffer descriptor) and a z-march stays inside a few memory pages.
//   Periodic runs store no z ghost rows (nzs = nz; the z boundary is an index
//   wrap inside the stencil loads).  Runs with an open z boundary store
//   zoff = gz ghost rows at each z end (nzs = nz + 2 gz), refilled before
//   every stage.
//   Velocity moments: 12 doubles {A0..A3, B0..B3, C0..C3} per site with
//   A = sum w u (bemr-emr), B = sum w u (bemi+emi), C = sum w u
//   (bee-bmm-ee+mm), u = (1, vx, vy, vz).  The stage kernels write partial
//   records [nx*nz sites][nch chunks][12] (each summed over the lpr
//   directions of one chunk) plus one spare x row of records; k_moment_totals
//   adds them in chunk order into totals [nx*nz sites][16], which the next
//   stage reads.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstdint>
#include <vector>

struct nuosc_sim;   // the handle (nuosc_capi.cpp)

namespace nuosc {

constexpr int kNVar = 8;         // fields of the 2-flavour state
constexpr int kNVar3 = 18;       // fields of the 3-flavour state
constexpr int kNMom3 = 32;       // 3-flavour moment doubles per site (8 vectors of 4)
constexpr int kWave = 64;
constexpr int kMomStride = 16;   // doubles per per-site moment total record (12 used)
constexpr int kMomRec = 12;      // doubles per (site, chunk) partial moment record
constexpr int kNMom = 12;
constexpr int kNAna = 12;        // analysis accumulators (see AnaIdx)
constexpr int kXg = 3;           // stored periodic ghost x rows at each end (the widest stencil's radius)

// advection schemes of the right-hand side (the reference's compile-time
// variants, nuosc_class.cpp:127-183; kAdvCfd4 is its default)
enum AdvScheme { kAdvCfd4 = 0, kAdvCfd2, kAdvUpwind2, kAdvLopsided, kAdvMuscl, kAdvOff };

enum AnaIdx {
    ANA_SURV = 0, ANA_SURVB, ANA_AVGP, ANA_AVGPB, ANA_AM01, ANA_AM02, ANA_AM03,
    ANA_NOR, ANA_NORB, ANA_REX2, ANA_MAXDP, ANA_BAD
};

struct Geometry {
    int nx, nz, nv, nvp;
    int lpr;               // directions per chunk (dirs_per_chunk)
    int rows;              // x rows per stage-kernel wavefront (rows_per_wave)
    int wpb;               // wavefronts per stage-kernel block (waves_per_block)
    int nch;               // direction chunks per site = ceil(nv / lpr)
    int nxs;               // stored x rows: nx + 2 kXg + 1 (ghost rows and the spare row)
    long long srec;        // elements between consecutive site records of one copy (nf nvp)
    long long xrow;        // elements between stored x rows of one copy: nzs * srec
    double twelve_dx, twelve_dz;   // 12*dx, 12*dz (advection denominators)
    double ko_x, ko_z;     // -ko/dx/16, -ko/dz/16
    double dx, dz;
    int ko5;               // 5th-order Kreiss-Oliger (radius 3) instead of the 3rd-order default
    int adv;               // AdvScheme
    const double *vx, *vy, *vz, *vw;   // device, length nv
    int bc;                // 0 periodic, 1 open in z
    int nzs, zoff;         // stored z rows per x row, stored row of interior z = 0
    int nf;                // fields per site record: 8 (2 flavours) or 18 (3 flavours)
    int nox;               // one x row (nx = 1) and the default discretisation: the stage
                           // kernel's NOX form (x neighbours are the point itself)
};

// k mod n in 0 .. n-1
__host__ __device__ inline int wrapi(int k, int n) {
    k %= n;
    return k < 0 ? k + n : k;
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Return the logical block index so that each XCD works on one contiguous
// range of logical blocks (a bijection for any grid size).
__host__ __device__ inline int xcd_remap(int b, int nb) {
    const int per = nb / 8, rem = nb % 8;
    const int x = b % 8, loc = b / 8;
    return (x < rem) ? x * (per + 1) + loc : rem * (per + 1) + (x - rem) * per + loc;
}

// element offset of field f, stored z row js, direction 0 of interior x row i
// (i = -kXg .. nx + kXg; nx + kXg is the spare row)
__host__ __device__ inline long long elem_off(const Geometry& g, long long i, long long js, int f) {
    return (i + kXg) * g.xrow + js * g.srec + (long long)f * g.nvp;
}
// element offset of direction 0, field 0 of interior site (i * nz + j)
__host__ __device__ inline long long site_off(const Geometry& g, long long site) {
    return elem_off(g, site / g.nz, site % g.nz + g.zoff, 0);
}
// elements spanned by one state copy (from its first element)
inline long long copy_elems(const Geometry& g) { return (long long)g.nxs * g.xrow; }

// Inflow data of the open z boundary: the ipt-0 pulse
// a = eps0 exp(-(z - vz t - zc)^2 / (2 sigma^2)) (on = 0: a = 0).
struct Inflow {
    int on;
    double eps0, sigma, zc, z0, dz;
};

struct Physics {
    double mu, pmo, ct, st;
    // 3-flavour vacuum Hamiltonian {hee, hmm, htt, hemr, hemi, hmtr, hmti,
    // hter, htei} (nuosc_class.cpp:543-569)
    double h3[9];
};

// One RK4 stage over the whole grid (k = L(yin)).
//   stage 0: yout = k                              (RHS only; test primitive)
//   stage 1: yout = yin + cy*k                     (yin is the state u)
//   stage 2: w = (yin + u) + cw*k,  yout = u + cy*k
//   stage 3: yout = u + cy*k
//   stage 4: yout = (w + yin)/3 + cw*k            (yout is the new state)
// With cy = dt/2, (dt, dt/2), dt and cw = dt/6 this is classic RK4 in 12
// passes over the state (w holds A = Y2 + 2 Y3 - u).
// SSP-RK3 (k = L(yin)):
//   stage 5: yout = yin + cy*k
//   stage 6/7: yout = ca*u + cb*(yin + cy*k)        (stage 7 may write u in place)
// mtot holds the moment totals of yin; mpart (stages 1-4) receives the partial
// moments of yout.
struct StagePlan {
    int zlen, nseg, nquads, ntask, nblocks;   // z-segment length and count, x-row quads, waves, blocks
};

// stencil radius of a discretisation: 3 for the 5th-order KO or the lopsided
// advection, else 2
inline int scheme_radius(int ko5, int adv) { return (ko5 || adv == kAdvLopsided) ? 3 : 2; }

// x-rows per wavefront of the stage kernel (1 row x 64 directions in both
// precisions)
int rows_per_wave(int esize);
// directions per chunk (64)
int dirs_per_chunk(int esize);
// wavefronts per stage-kernel block (4)
int waves_per_block(int esize);

StagePlan plan_stage(const Geometry& g, int num_cus);
// the same decomposition with a given z-segment length
StagePlan plan_stage_zlen(const Geometry& g, int zlen);
// The stage kernel addresses the stored x rows of one wavefront (rows + 5 of
// them) and the moment records through buffer descriptors: each range must
// span fewer than 4 GiB.
bool buffer_addressable(const Geometry& g, int esize);
// The same for the 3-flavour z march (boxes with nx > 1; one-row grids take
// a kernel with 64-bit addressing).
bool flavor3_addressable(const Geometry& g, int esize);

template <typename T>
hipError_t launch_stage(int stage, const Geometry& g, const Physics& ph, const StagePlan& pl,
                        const T* yin, const T* u, T* w, T* yout, const double* mtot, double* mpart,
                        double cw, double cy, hipStream_t s, double ca = 0.0, double cb = 0.0);

// periodic x ghost rows of a state copy: stored rows 0 .. kXg-1 <- interior
// rows nx-kXg .. nx-1 and rows nx+kXg .. nx+2kXg-1 <- interior 0 .. kXg-1
// (whole stored rows, z ghost rows included).  Run on every stage input
// before the stage.
template <typename T>
hipError_t launch_xghosts(const Geometry& g, T* y, hipStream_t s);

// partial records -> per-site totals (compensated sums for the FP32 form)
hipError_t launch_moment_totals(const Geometry& g, bool compensated, const double* part, double* tot,
                                hipStream_t s);

// partial moment records of a state (after init / upload)
template <typename T>
hipError_t launch_moments(const Geometry& g, const T* y, double* mpart, hipStream_t s);

// Initial data.  ipt 0/2/3: per-z perturbation tables (tmpr, p3o) computed on
// the host; ipt 5: seeded complex perturbation generated per point.
template <typename T>
hipError_t launch_init(const Geometry& g, int ipt, const double* G0, const double* G0b,
                       const double* tmpr_z, const double* p3o_z, double eps0, uint64_t seed,
                       T* y, hipStream_t s);

// Diagnostics: per-block partials then one fixed-order final reduction into
// out[kNAna] (device).  part must hold ana_blocks() * kNAna * 2 doubles.
int ana_blocks(const Geometry& g);
template <typename T>
hipError_t launch_analysis(const Geometry& g, const T* y, const double* G0, const double* G0b,
                           double* part, double* out, hipStream_t s);

// Reference-layout copy of field f of state y with its ghost band filled by
// the periodic rule (corners zero), as FP64: dst[(nx+2gx)*(nz+2gz)*nv].
template <typename T>
hipError_t launch_ghosted(const Geometry& g, int gx, int gz, const T* y, int f, double* dst,
                          hipStream_t s);

// Interior of field f of state y <-> FP64 dense [nx][nz][nv].
template <typename T>
hipError_t launch_pack(const Geometry& g, const T* y, int f, double* dst, hipStream_t s);
template <typename T>
hipError_t launch_unpack(const Geometry& g, const double* src, T* y, int f, hipStream_t s);

// P3 = (ee-mm)/G0 at x indices xi[sx] and v indices vi[sv]: dst[sx][nz][sv].
template <typename T>
hipError_t launch_p3_slice(const Geometry& g, const T* y, const double* G0, const int* xi, int sx,
                           const int* vi, int sv, double* dst, hipStream_t s);

// Batched runs: every member run advances one RK4 step per batch_step call
// with one launch per stage for all members (plus one moment-total launch per
// stage).  Members keep their own buffers, plans and physics; all share the
// precision (esize) of the tables.
struct BatchMember {
    Geometry g;
    Physics ph;
    StagePlan pl;
    void *u, *w, *ya, *yb;          // state copies
    double *mpart, *mu, *ma, *mb;   // partial records; totals of u, yA, yB
    double dt;
};
// Launch forms of the batched stage kernel: an ordinary grid of one block
// per logical block, or a persistent grid of resident blocks drawing logical
// blocks from per-XCD work queues.
constexpr int kBatchGrid = 0, kBatchPersistent = 1;
constexpr int kBatchCounters = 32;   // 8 XCD counters for each of the 4 stages

struct BatchTables {
    void* args[4] = {nullptr, nullptr, nullptr, nullptr};  // per-stage StageArgs arrays
    void* tot[4] = {nullptr, nullptr, nullptr, nullptr};   // per-stage moment-total arrays
    void* xg[4] = {nullptr, nullptr, nullptr, nullptr};    // per-stage x-ghost refresh arrays
    int* bmap = nullptr;   // logical block -> (member, block of member)
    int* tmap = nullptr;
    int* xmap = nullptr;
    int* ctr = nullptr;    // work-queue counters of the persistent launch
    int nblocks = 0, tblocks = 0, xblocks = 0, esize = 8;
    int launch = kBatchGrid, pgrid = 0;   // launch form; resident blocks of the persistent grid
};
hipError_t batch_build(BatchTables& t, const std::vector<BatchMember>& members, int esize, int launch,
                       hipStream_t s);
void batch_free(BatchTables& t);
hipError_t batch_step(const BatchTables& t, hipStream_t s);
// one stage kernel of every member (no ghost refresh, no moment totals):
// autotuner timing
hipError_t batch_stage(const BatchTables& t, int stage, hipStream_t s);

// Open z boundary: fill the zoff ghost rows at both z ends of state y for
// time t = *t_dev + cdt (inflow directions: the Inflow pulse state;
// outflow directions: quadratic extrapolation of the three edge cells).
template <typename T>
hipError_t launch_open_ghosts(const Geometry& g, T* y, const double* G0, const double* G0b, const Inflow& in,
                              const double* t_dev, double cdt, hipStream_t s);
// *t_dev += dt (device copy of the physical time, read by the boundary fill)
hipError_t launch_advance_time(double* t_dev, double dt, hipStream_t s);

// Three flavours (nf = 18; the reference's default discretisation, KO 3 and
// central 4th-order advection).  Moments: per site 32 doubles, the vectors
// {emRm, emIp, eemm, mtRm, mtIp, teRm, teIp, mmtt} x (1, vx, vy, vz)
// (nuosc_class.cpp:28-82), summed over all directions of y in a fixed order.
template <typename T>
hipError_t launch_moments3(const Geometry& g, const T* y, double* mom, hipStream_t s);
// One stage over the whole grid with the moment totals mom of yin; the stage
// numbers and coefficients of launch_stage.  Stages 1-7 also write the
// partial moments of yout to part ([site][ceil(nv/64)][32] doubles), which
// launch_moment_totals3 turns into the totals of yout.
template <typename T>
hipError_t launch_stage3(int stage, const Geometry& g, const Physics& ph, const T* yin, const T* u, T* w,
                         T* yout, const double* mom, double* part, double cw, double cy, hipStream_t s,
                         double ca = 0.0, double cb = 0.0);
hipError_t launch_moment_totals3(const Geometry& g, const double* part, double* mom, hipStream_t s);

// What an entry point implemented beside the kernels needs of a handle: its
// geometry, current state u, G0 | G0b, value size, device and stream.
struct SimView {
    Geometry geo;
    const void* state;
    const double* g0;
    int esize;
    int device;
    hipStream_t stream;
};
SimView sim_view(const nuosc_sim* s);
// records msg as the handle's last error and returns code
int sim_fail(nuosc_sim* s, int code, const char* msg);

// Renormalisation of the state to |P| = 1 (the reference's formula).
template <typename T>
hipError_t launch_renorm(const Geometry& g, T* y, const double* G0, const double* G0b,
                         hipStream_t s);

}  // namespace nuosc
