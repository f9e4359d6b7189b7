/*
 * nuosc_hip.h -- C ABI of libnuosc_hip.so, the MI355X (gfx950) implementation
 * of the nuosc RK4 time-stepper core.
 *
 * The reference exposes this path through the C++ class NuOsc
 * (nuosc_class.h:99-191) and the ./nuosc command line (nuosc.cpp:17-261); it
 * has no plugin or FFI layer.  This header is the plain-C boundary beneath a
 * NuOsc-compatible host class (nuosc_amd/csrc/NuOsc.h) and the ctypes binding
 * (nuosc_amd/__init__.py).  Every entry point below cites the reference
 * member it replaces.
 *
 * Conventions
 *   - A handle (nuosc_sim*) owns all device memory; it is driven by one host
 *     thread and uses one HIP stream.  Distinct handles are independent.
 *   - Every function returns 0 on success or a nonzero NUOSC_E* code; the
 *     message is available from nuosc_last_error(handle) (or
 *     nuosc_last_error(NULL) for errors of nuosc_create).  No C++ exception
 *     crosses this boundary.
 *   - Host arrays are FP64.  Field order is the reference's
 *     (getAllFields, nuosc_class.h:58-65): 0 ee, 1 mm (=xx), 2 emr, 3 emi, 4 bee, 5 bmm,
 *     6 bemr, 7 bemi.  The "interior" layout is [nx][nz][nv] (v fastest); the
 *     "ghosted" layout is the reference's lpts layout
 *     ((i+gx)*(nz+2gz)+j+gz)*nv+v (CartGrid.h:91).
 */
#ifndef NUOSC_HIP_H
#define NUOSC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NUOSC_ABI_VERSION 5

/* status codes */
#define NUOSC_OK 0
#define NUOSC_EINVAL 1      /* bad argument or configuration            */
#define NUOSC_EHIP 2        /* a HIP runtime call failed                */
#define NUOSC_ENOMEM 3      /* device or host allocation failed         */
#define NUOSC_EUNSTABLE 4   /* analysis: maxdP >= 10 or non-finite state */
#define NUOSC_ENODEV 5      /* no usable GPU                            */
#define NUOSC_EAGAIN 6      /* server: no result ready within the wait  */

#define NUOSC_PREC_FP64 64
#define NUOSC_PREC_FP32 32

#define NUOSC_VGRID_VORONOI 0 /* icosahedron.h weights (reference)          */
#define NUOSC_VGRID_EQUAL 1   /* same directions, weights 4*pi/N            */

#define NUOSC_BC_PERIODIC 0 /* nuosc_boundary.cpp:4-52                                  */
#define NUOSC_BC_OPEN_Z 1   /* open z ends, x periodic.  The reference's open boundary,
                             * updateInjetOpenBoundary (nuosc_boundary.cpp:54-57), is a
                             * stub; this one has no reference counterpart.  Before every
                             * RK stage the gz ghost rows at each z end are refilled for
                             * the stage time t: a direction entering the box there
                             * (vz > 0 at z0, vz < 0 at z1) gets the ipt-0 state of
                             * amplitude a = eps0 exp(-(z - vz t - zc)^2 / (2 sigma^2))
                             * (eps0, sigma of the last fill_init if it used ipt 0, else
                             * a = 0: the unperturbed state); a leaving direction gets the
                             * quadratic extrapolation of the three edge cells.  Needs
                             * nz >= 3 and gz in {2, 3}; not batchable. */

/* Discretisation of calRHS: the reference's compile-time variants
 * (nuosc_class.cpp:112-183, selected there by KO_ORD_3 and ADVEC_*), chosen
 * per handle here.  Kreiss-Oliger order 3 (KO_ORD_3, the default: -ko/(16h)
 * delta^4) or 5 (-ko/(64h) delta^6, as written in the reference, whose sign
 * makes it amplify rather than damp the highest wavenumbers); advection: */
#define NUOSC_ADV_CFD4 0      /* 4th-order central (default) */
#define NUOSC_ADV_CFD2 1      /* ADVEC_CFD2O: 2nd-order central */
#define NUOSC_ADV_UPWIND2 2   /* ADVEC_UPWIND2O: 2nd-order upwind */
#define NUOSC_ADV_LOPSIDED 3  /* ADVEC_LOPESIDE: 4th-order lopsided */
#define NUOSC_ADV_MUSCL 4     /* ADVEC_MUSCL: flux limited (superbee) */
#define NUOSC_ADV_OFF 5       /* ADVEC_OFF */

#define NUOSC_NVAR2 8         /* fields of a 2-flavour state */
#define NUOSC_NVAR3 18        /* fields of a 3-flavour state */
/* KO order 5 and the lopsided scheme have stencil radius 3: with the open z
 * boundary they need gz = 3.  Handles with a non-default discretisation are
 * not batchable. */

typedef struct nuosc_sim nuosc_sim;

/* Run configuration.  Replaces the NuOsc constructor arguments
 * NuOsc(px,pz,nv,nphi,gx,gz,x0,x1,z0,z1,dx,dz,CFL,ko) (nuosc_class.h:131-133,
 * nuosc_class.cpp:454-466) plus set_mu/set_pmo/set_renorm
 * (nuosc_class.h:148-159).  px,pz (MPI) and nphi (polar grid only) have no
 * counterpart here. */
typedef struct nuosc_config {
    int nn;          /* icosahedron refinement (--nv): N = 10*nn^2+2 directions */
    int gx, gz;      /* ghost widths of the reference layout (2 by default) */
    double x0, x1, z0, z1, dx, dz;
    double cfl;      /* dt = dz*cfl (nuosc_class.cpp:464) */
    double ko;       /* Kreiss-Oliger coefficient */
    double mu;       /* self-interaction strength */
    double pmo;      /* vacuum term: 1/-1 mass ordering, 0 off */
    int renorm;      /* renormalise |P| after each step */
    int vgrid;       /* NUOSC_VGRID_* */
    int precision;   /* NUOSC_PREC_FP64 or NUOSC_PREC_FP32 */
    int bc;          /* NUOSC_BC_PERIODIC or NUOSC_BC_OPEN_Z */
    int device;      /* HIP device ordinal */
    int use_graph;   /* capture one step as a hipGraph and replay it */
    double zc;       /* centre of the ipt-0 profile eps0 exp(-(z-zc)^2/(2 sigma^2));
                      * the reference's is 0 (nuosc_init.cpp:4, 98) */
    int ko_order;    /* Kreiss-Oliger order: 3 (or 0: the default) or 5 */
    int adv;         /* NUOSC_ADV_* (0: the default central scheme) */
    int nflavor;     /* 2 (or 0: the default) or 3; see the field order below */
} nuosc_config;

/* Fields.  2 flavours (the reference's default build): 8 fields, index
 * 0..7 = ee, mm, emr, emi, bee, bmm, bemr, bemi.  3 flavours (its
 * compile-time N_FLAVOR_3, common.h:10-16; nuosc_class.h:17-67): 18 fields,
 * the same 8 first, then 8..17 = tt, mtr, mti, ter, tei, btt, bmtr, bmti,
 * bter, btei (the reference's getAllFields order is ee, mm, tt, emr, emi,
 * mtr, mti, ter, tei, bee, bmm, btt, bemr, bemi, bmtr, bmti, bter, btei).
 * The 3-flavour right-hand side is nuosc_class.cpp:28-82, 189-277 as
 * written, with the vacuum Hamiltonian of :543-569 (mixing angles given in
 * degrees/180 and used as radians, delta_cp = 0).  fillInitValue writes the
 * 8 two-flavour fields and zeroes the tau fields; analysis() and
 * renormalize() act on the two-flavour fields, as in the reference.
 * 3-flavour handles take the default discretisation only and are not
 * batchable. */

/* Diagnostics of one analysis() call (nuosc_ana.cpp:34-137).  The first nine
 * are the analysis.dat columns; the rest are additions. */
typedef struct nuosc_analysis_t {
    double phy_time;
    double maxdP;    /* max |1-|P|| over nu and nubar                   */
    double surv;     /* Pee:    sum w ee / sum w G0                     */
    double survb;    /* Peebar: sum w bee / sum w G0b                   */
    double avgP;     /* sum w G0 |1-|P|| / sum w G0                     */
    double avgPb;
    double aM0;      /* |M0| * ds_L                                     */
    double Lex;      /* M0_3 * ds_L                                     */
    double ELNe;
    /* additions (not written to analysis.dat) */
    double rho_ex2;  /* sum w (emr^2 + emi^2) * ds_L: |rho_ex| growth   */
    double nonfinite;/* number of points with a non-finite field value  */
    int status;      /* 0, or NUOSC_EUNSTABLE                            */
} nuosc_analysis_t;

/* Library identity. */
int nuosc_abi_version(void);
const char* nuosc_build_info(void);

/* Host-only (no GPU call): the velocity grid a handle would use, as
 * gen_v2d_icosahedron (CartGrid.cpp:46-61) builds it.  mode is NUOSC_VGRID_*.
 * Arrays must hold 10*nn^2+2 doubles; returns that count, or -1. */
int nuosc_vgrid(int nn, int mode, double* vx, double* vy, double* vz, double* vw);

/* NuOsc::NuOsc (nuosc_class.cpp:454-534): geometry, velocity grid,
 * allocation.  The state is zero until nuosc_fill_init or an upload.
 * NUOSC_EINVAL also for grids the stage kernels cannot address: their 32-bit
 * buffer offsets span one block's stored x rows (12 rows of nz * nf * nvp
 * values; 2 flavours: also the moment records) and must stay under 4 GiB
 * (one x row of 3 flavours is not limited). */
int nuosc_create(const nuosc_config* cfg, nuosc_sim** out);
/* NuOsc::~NuOsc (nuosc_class.h:136-146) */
int nuosc_destroy(nuosc_sim* sim);
const char* nuosc_last_error(const nuosc_sim* sim);

/* NuOsc::set_mu / set_pmo / set_renorm (nuosc_class.h:148-159) */
int nuosc_set_mu(nuosc_sim* sim, double mu);
int nuosc_set_pmo(nuosc_sim* sim, double pmo);
int nuosc_set_renorm(nuosc_sim* sim, int renorm);

/* Grid queries: CartGrid members nx, nz, nv, lpts (CartGrid.h:27-34,85-89)
 * and NuOsc members dt, ds_L, phy_time, n_nue0 (nuosc_class.h:103-114). */
int nuosc_get_dims(const nuosc_sim* sim, int* nx, int* nz, int* nv, long long* lpts);
int nuosc_get_scalars(const nuosc_sim* sim, double* dt, double* ds_L, double* phy_time,
                      double* n_nue0_2);
/* Sets phy_time (also the time the open z boundary data is evaluated at). */
int nuosc_set_phy_time(nuosc_sim* sim, double phy_time);
/* CartGrid vx, vy, vz, vw (length nv) and X (nx), Z (nz) (CartGrid.h:29-32) */
int nuosc_get_vgrid(const nuosc_sim* sim, double* vx, double* vy, double* vz, double* vw);
int nuosc_get_coords(const nuosc_sim* sim, double* X, double* Z);
/* G0, G0b per direction (the reference stores them per point,
 * nuosc_class.h:113). */
int nuosc_get_g0(const nuosc_sim* sim, double* G0, double* G0b);
/* Device bytes held by the handle. */
int nuosc_get_device_bytes(const nuosc_sim* sim, long long* bytes);

/* NuOsc::fillInitValue(ipt,alpha,eps0,sigma,lnue,lnueb,lnuex,lnuebx)
 * (nuosc_init.cpp:17-137) for ipt 0, 2, 3; ipt 5 adds a seeded complex
 * perturbation rho_ex = eps0*(ur + i ui)*G0/2 with (ur, ui) uniform in [-1,1)
 * from a counter-based generator keyed by (seed, interior index).  ipt 1 and
 * 4 use the C library rand() in the reference and are rejected. */
int nuosc_fill_init(nuosc_sim* sim, int ipt, double alpha, double eps0, double sigma, double lnue,
                    double lnueb, double lnuex, double lnuebx, uint64_t seed);

/* NuOsc::step_rk4 (nuosc_class.cpp:374-427), nsteps times; phy_time += dt
 * per step.  Asynchronous on the handle's stream. */
int nuosc_step_rk4(nuosc_sim* sim, int nsteps);
/* NuOsc::step_srk3 (nuosc_class.cpp:429-452), the 3-stage SSP Runge-Kutta
 * step, nsteps times; 8 state passes per step instead of RK4's 16.  Unlike
 * the reference's step_srk3, which refreshes no ghost cells, every stage
 * input gets the boundary fill (stage times t, t+dt, t+dt/2). */
int nuosc_step_srk3(nuosc_sim* sim, int nsteps);
int nuosc_synchronize(nuosc_sim* sim);

/* NuOsc::calRHS(out, in) (nuosc_class.cpp:11-307) with the boundary fill of
 * its input (periodic: nuosc_boundary.cpp:4-52; open z: the NUOSC_BC_OPEN_Z
 * rule at the current phy_time): in/out are host arrays of
 * nvar*nx*nz*nv doubles in the interior layout (nvar = 8 or 18 fields).  Test primitive; uses the
 * handle's scratch state buffers (the state itself is left untouched). */
int nuosc_calrhs(nuosc_sim* sim, const double* in, double* out);

/* NuOsc::analysis (nuosc_ana.cpp:34-137) on the current state. */
int nuosc_analysis(nuosc_sim* sim, nuosc_analysis_t* out);

/* NuOsc::renormalize (nuosc_ana.cpp:140-166) on the current state. */
int nuosc_renormalize(nuosc_sim* sim);

/* State transfer (replaces direct access to v_stat->ee etc.); field is a
 * field index of the order above.
 * download_field: one field, interior layout (nx*nz*nv doubles).
 * download_field_ghosted: one field in the reference's ghosted layout with
 *   the ghost band filled by updatePeriodicBoundary's rule
 *   (nuosc_boundary.cpp:4-52), or for NUOSC_BC_OPEN_Z by the open rule at the
 *   current phy_time in z, and zero corners ((nx+2gx)(nz+2gz)nv doubles).
 * upload_field: one field from the interior layout.  After uploading all the
 *   fields call nuosc_state_updated to refresh the velocity moments. */
int nuosc_download_field(nuosc_sim* sim, int field, double* host);
int nuosc_download_field_ghosted(nuosc_sim* sim, int field, double* host);
int nuosc_upload_field(nuosc_sim* sim, int field, const double* host);
int nuosc_state_updated(nuosc_sim* sim);

/* P3 snapshot slice (nuosc_snapshot.cpp:82-122 with var = P3, refreshed as
 * in eval_conserved, nuosc_ana.cpp:14): dst[sx][nz][sv] with
 * P3 = (ee-mm)/G0 at x indices xidx[sx] and direction indices vidx[sv]. */
int nuosc_p3_slice(nuosc_sim* sim, const int* xidx, int sx, const int* vidx, int sv, double* dst);
/* No counterpart in the reference.  The same slice for callers that cannot
 * allocate device memory (serving loops, code between graph replays,
 * harnesses that forbid allocation): nuosc_p3_slice may hipMalloc a larger
 * staging buffer; this variant stages the index lists and the output in a
 * fixed 16 MiB device array that each GPU holds for the whole process, and
 * calls no allocator.  Limits: sx + sv <= 16384 and sx * nz * sv <= 2088960,
 * else NUOSC_EINVAL.  Calls on one GPU share the array: calls of this
 * process serialise on a per-device lock (each call is synchronous). */
int nuosc_p3_slice_noalloc(nuosc_sim* sim, const int* xidx, int sx, const int* vidx, int sv, double* dst);

/* Wall-clock of the device work of the last nuosc_step_rk4 call (ms),
 * from HIP events on the handle's stream (synchronises). */
int nuosc_last_step_ms(nuosc_sim* sim, double* ms);

/* Measurement helper: run nsteps RK4 steps eagerly with a pair of HIP events
 * around each of the four stage kernels (on the handle's stream) and return
 * the average duration of each stage kernel in stage_ms[4] (ms).  Advances
 * the state and phy_time like nuosc_step_rk4. */
int nuosc_profile_stages(nuosc_sim* sim, int nsteps, double* stage_ms);

/* Stage-kernel decomposition (no reference counterpart).  Each wavefront
 * marches along z over a segment of zlen sites; nuosc_autotune times RK4
 * stages 1-3 (with their boundary fills and moment totals) on this handle's
 * own grid for zlen in {8, ..., 256}, the candidates in turn for trials
 * rounds after a warm-up round, keeps each candidate's fastest time and
 * selects the best.  Call it at a step boundary: it overwrites only the RK4
 * scratch copies, never the state.  Results are cached per (nx, nz, nv,
 * precision, bc, ko_order, adv, one-row form) for the process.  3-flavour handles have no segment length:
 * autotune returns NUOSC_OK with *best_zlen = 0 and set_zlen has no effect. */
int nuosc_get_plan(const nuosc_sim* sim, int* zlen, int* waves);
int nuosc_set_zlen(nuosc_sim* sim, int zlen);
int nuosc_autotune(nuosc_sim* sim, int trials, int* best_zlen, double* best_ms);

/* ------------------------------------------------------------------------
 * Batched runs (no reference counterpart: the reference runs one NuOsc per
 * process).  A batch advances many independent handles -- different grids,
 * direction counts and physics -- with one launch per RK4 stage for all of
 * them.  Members keep their own state and remain usable through the
 * single-run functions between batch steps; a member's results are bitwise
 * the same stepped alone or in a batch.  All members share the batch's
 * device and precision; renormalising members are not supported.  A handle
 * belongs to at most one batch.  Changing a member's physics or
 * decomposition (set_mu, set_pmo, set_zlen, autotune) takes effect at the
 * next batch step; destroying a member removes it from its batch first.
 * ------------------------------------------------------------------------ */
typedef struct nuosc_batch nuosc_batch;
int nuosc_batch_create(int device, int precision, nuosc_batch** out);
int nuosc_batch_destroy(nuosc_batch* batch);
int nuosc_batch_add(nuosc_batch* batch, nuosc_sim* sim);
int nuosc_batch_remove(nuosc_batch* batch, nuosc_sim* sim);
int nuosc_batch_size(const nuosc_batch* batch);
/* nsteps RK4 steps of every member (phy_time += dt per step each);
 * returns after the device work completed. */
int nuosc_batch_step(nuosc_batch* batch, int nsteps);
const char* nuosc_batch_last_error(const nuosc_batch* batch);
/* z-segment length of the batched stage kernels, shared by every member
 * (default 64, capped at each member's nz; 0 = each member's own plan).  A
 * common length gives every block of a batched launch the same work.
 * batch_autotune times the batched stage-1 kernel for zlen in {16, ..., 256}
 * and both launch forms (trials launches each, the fastest kept; only the
 * RK4 scratch copies are written) and keeps the best pair (the launch form
 * it chose: nuosc_batch_get_launch).  Results do not depend on zlen or the
 * launch form. */
int nuosc_batch_set_zlen(nuosc_batch* batch, int zlen);
int nuosc_batch_autotune(nuosc_batch* batch, int trials, int* best_zlen, double* best_ms);
/* Launch form of the batched stage kernels: NUOSC_BATCH_GRID (default: one
 * block per logical block of every member) or NUOSC_BATCH_PERSISTENT (one
 * grid of as many blocks as the GPU holds at once, drawing logical blocks
 * from per-XCD work queues until they are empty).  get_launch returns the
 * form, or -1 for a NULL batch. */
#define NUOSC_BATCH_GRID 0
#define NUOSC_BATCH_PERSISTENT 1
int nuosc_batch_set_launch(nuosc_batch* batch, int launch);
int nuosc_batch_get_launch(const nuosc_batch* batch);

/* Host-only: device bytes a handle created from cfg would hold (admission
 * control for the serving loop). */
int nuosc_estimate_device_bytes(const nuosc_config* cfg, long long* bytes);

/* ------------------------------------------------------------------------
 * Run server (no reference counterpart).  Parameter-scan runs are submitted
 * by any host thread; one dispatcher thread admits each queued run whose
 * device footprint fits beside the resident runs' within the HBM budget
 * (scanning the queue in arrival order, so a small run is not held behind a
 * large one; once the oldest queued run has waited longer than the age
 * bound, 30 s by default, no younger run is admitted before it), advances all
 * resident runs together through a batch (continuous batching: runs join and
 * leave between ticks of at most max_tick_steps steps) and records each run's
 * (t, Pee, Peebar, maxdP, |M0|) series every ana_every steps.  submit
 * refuses (NUOSC_EINVAL) runs a batch cannot take (open z, a non-default
 * discretisation, 3 flavours, renormalisation) and runs whose footprint
 * alone exceeds the budget.
 * ------------------------------------------------------------------------ */
typedef struct nuosc_server nuosc_server;

typedef struct nuosc_request {
    nuosc_config cfg;   /* geometry and physics (device/precision from the server) */
    int ipt;
    double alpha, eps0, sigma, lnue, lnueb, lnuex, lnuebx;
    uint64_t seed;
    int nsteps;
    int ana_every;
} nuosc_request;

typedef struct nuosc_result {
    long long id;
    int status;          /* 0 or the NUOSC_E* code that stopped the run */
    int nsteps;          /* steps done */
    double t_submit, t_admit, t_done;   /* seconds on the server's clock */
    long long device_bytes;
    int n_samples;       /* series rows of 5 doubles: t, Pee, Peebar, maxdP, |M0| */
} nuosc_result;

typedef struct nuosc_server_stats_t {
    int queued, active;
    long long finished;
    long long resident_bytes;
    double updates;      /* site*dir RK4 updates done so far */
    double busy_s;       /* dispatcher time inside batch steps */
    double now_s;
} nuosc_server_stats_t;

int nuosc_server_create(int device, int precision, long long hbm_budget_bytes, int max_tick_steps,
                        nuosc_server** out);
int nuosc_server_submit(nuosc_server* srv, const nuosc_request* req, long long* id);
/* id < 0: any finished run.  Waits up to timeout_s; NUOSC_EAGAIN if none.
 * series (optional) receives min(n_samples, max_samples) rows. */
int nuosc_server_take_result(nuosc_server* srv, long long id, double timeout_s, nuosc_result* info,
                             double* series, int max_samples);
int nuosc_server_stats(nuosc_server* srv, nuosc_server_stats_t* stats);
/* Seconds the oldest queued run may be passed over by younger runs that fit
 * (>= 0; 0 = strict arrival order). */
int nuosc_server_set_age_bound(nuosc_server* srv, double seconds);
/* Stops the dispatcher: queued and resident runs are cancelled. */
int nuosc_server_destroy(nuosc_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* NUOSC_HIP_H */
