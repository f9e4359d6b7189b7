/*
 * nuosc_oracle.h -- CPU restatement of the nuosc RK4 time-stepper core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path (libnuosc_hip.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * It restates, in plain C99 + OpenMP, the reference's algorithm with the
 * reference's own storage layout (8 SoA fields with a ghost band, index
 * ((i+gx)*(nz+2gz)+j+gz)*nv+v, CartGrid.h:91) and the reference's operation
 * order, so that its results reproduce the reference's arithmetic:
 *
 *   orc_vgrid_icosahedron   icosahedron.h:36-171, Point.h:64-69, CartGrid.cpp:46-61
 *   orc_create (geometry)   CartGrid.cpp:191-208, nuosc_class.cpp:454-466
 *   orc_fill_init           nuosc_init.cpp:4-137 (ipt 0/2/3) + seeded ipt 5
 *   orc_periodic_ghosts     nuosc_boundary.cpp:4-52
 *   orc_open_ghosts         open z boundary (the reference's
 *                           updateInjetOpenBoundary, nuosc_boundary.cpp:54-57,
 *                           is a stub): inflow ghosts carry the free-streaming
 *                           ipt-0 pulse, outflow ghosts are quadratic
 *                           extrapolations of the three edge cells
 *   orc_calrhs              nuosc_class.cpp:11-307 (2 flavours, KO_ORD_3,
 *                           central 4th-order advection, vacuum on; the
 *                           discretisation variants of :112-183; 3 flavours,
 *                           N_FLAVOR_3, :189-277 with the Hamiltonian of
 *                           :543-569)
 *   orc_step_rk4            nuosc_class.cpp:312-351, 374-427
 *   orc_step_srk3           nuosc_class.cpp:353-372, 429-452 (with boundary
 *                           fills, which the reference's step_srk3 omits)
 *   orc_analysis            nuosc_ana.cpp:3-137
 *   orc_renormalize         nuosc_ana.cpp:140-166
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   - v-grid: bit-exact against the reference's own icosahedron.h compiled
 *     where it lies (oracle/_ref/ref_vgrid) and the committed fixtures made by
 *     it; closed forms N=10nn^2+2, |v|=1, sum w v = 0.
 *   - ghosts: brute-force modular-index check.
 *   - calRHS: an independent numpy formulation (2x2 complex density matrices,
 *     H = mu*sum_k w_k(1-v.v_k)(rho_k - conj(rhobar_k)) + pmo(st s1 - ct s3),
 *     rhs = -i[H,rho], np.roll finite differences); single-Fourier-mode symbol
 *     of ADV+KO.
 *   - SSP-RK3: exact amplification (1 + z + z^2/2 + z^3/6)^n of a Fourier
 *     mode.
 *   - RK4: exact amplification G(z)^n of a Fourier mode; vacuum precession
 *     polynomial; conservation of sum_xz(ee+mm) per direction.
 *   - init: sum_v w G0 = 1, sum_v w G0b = alpha, trace and purity identities.
 *   - analysis: an independent numpy formulation of the diagnostics plus the
 *     pure-state/conservation identities.
 *   - 3 flavours (the reference's N_FLAVOR_3 build stops at the preprocessor:
 *     `#ifdef 0`, nuosc_class.cpp:207, so no reference run exists): with the
 *     tau fields zero and pmo = 0 the 8 two-flavour rows equal the 2-flavour
 *     RHS bit for bit and the 10 tau rows are zero; the interaction part
 *     keeps Tr rho and Tr rho^2 of every direction (nu and nubar) for random
 *     states; hv equals an independent U diag(0, dm21, dm31) U^T - Tr/3 in
 *     the reference's angle convention.  The vacuum rows outside the e-mu
 *     block are parity unpinned (they are not a commutator as written: see
 *     orc_vacuum_h3).
 *   - open boundary (no reference counterpart): at mu=0 the exact solution is
 *     free streaming of the initial profile; the interior error converges at
 *     third order in dz (the order of the KO term) with inflow and outflow both
 *     active; the z-sum of the RHS equals the telescoped boundary flux.
 * Nothing else of the reference builds here: every other reference source
 * includes common.h, which hard-wires COSENU_MPI and therefore <mpi.h>,
 * which this image does not have.
 */
#ifndef NUOSC_ORACLE_H
#define NUOSC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fields per state: 8 for 2 flavours (ee, mm, emr, emi, bee, bmm, bemr,
 * bemi), 18 for 3 flavours (the same 8 first, then tt, mtr, mti, ter, tei,
 * btt, bmtr, bmti, bter, btei); arrays of field pointers hold ORC_NVAR */
#define ORC_NVAR 18
#define ORC_NVAR2 8
#define ORC_NVAR3 18

typedef struct orc_sim {
    /* geometry (CartGrid.cpp:191-208) */
    int nx, nz, nv, gx, gz;
    double x0, x1, z0, z1, dx, dz;
    double *X, *Z;
    double *vx, *vy, *vz, *vw;
    long lpts;
    /* physics (nuosc_class.h:99-126, nuosc_class.cpp:454-466) */
    double dt, ds_L, cfl, ko, mu, pmo, ct, st;
    int renorm;
    double phy_time;
    double n_nue0[2];
    double *G0, *G0b;          /* per direction (the reference stores them per point) */
    /* four RK4 state buffers, nvar (8 or 18) fields each (nuosc_class.h:108-109) */
    double *stat[ORC_NVAR], *rhs[ORC_NVAR], *pre[ORC_NVAR], *cor[ORC_NVAR];
    int analysis_lines;        /* analysis.dat line counter (first line precision quirk) */
    /* boundary: 0 periodic (reference), 1 open in z (x periodic) */
    int bc;
    double zc;                 /* centre of the ipt-0 profile (the reference uses 0) */
    int inflow_on;             /* inflow pulse active (fill_init with ipt 0) */
    double inflow_eps0, inflow_sigma;
    /* discretisation of calRHS (the reference's compile-time variants,
     * nuosc_class.cpp:112-183): Kreiss-Oliger order 3 (KO_ORD_3, default) or
     * 5, and the advection scheme ORC_ADV_* */
    int ko_order;
    int adv;
    /* flavours: nvar = 8 (2 flavours, default) or 18 (3 flavours, the
     * reference's compile-time N_FLAVOR_3, common.h:10-16) */
    int nvar;
    /* 3-flavour vacuum Hamiltonian {hee, hmm, htt, hemr, hemi, hmtr, hmti,
     * hter, htei} (nuosc_class.cpp:543-569) */
    double hv[9];
} orc_sim;

#define ORC_ADV_CFD4 0      /* 4th-order central (default) */
#define ORC_ADV_CFD2 1      /* ADVEC_CFD2O: 2nd-order central */
#define ORC_ADV_UPWIND2 2   /* ADVEC_UPWIND2O: 2nd-order upwind */
#define ORC_ADV_LOPSIDED 3  /* ADVEC_LOPESIDE: 4th-order lopsided */
#define ORC_ADV_MUSCL 4     /* ADVEC_MUSCL: flux-limited (superbee) */
#define ORC_ADV_OFF 5       /* ADVEC_OFF */

#define ORC_BC_PERIODIC 0
#define ORC_BC_OPEN_Z 1

/* Velocity grid.  Returns N = 10*nn*nn+2 and fills the four arrays (length N). */
int orc_vgrid_icosahedron(int nn, double* vx, double* vy, double* vz, double* vw);

/* Create.  vgrid_mode 0 = icosahedral Voronoi weights (reference), 1 = equal
 * weights 4*pi/N on the same directions.  Returns NULL on bad input. */
orc_sim* orc_create(int nn, int gx, int gz, double x0, double x1, double z0, double z1,
                    double dx, double dz, double cfl, double ko, int vgrid_mode);
void orc_destroy(orc_sim* s);
void orc_set_physics(orc_sim* s, double mu, double pmo, int renorm);
/* Boundary mode and ipt-0 profile centre; call before orc_fill_init.  The open
 * mode needs nz >= 3.  Returns 0, or -1 on bad input. */
int orc_set_boundary(orc_sim* s, int bc, double zc);
/* Discretisation (ko_order 3 or 5, adv ORC_ADV_*); the stencil radius (3 for
 * KO 5 and the lopsided scheme, else 2) must not exceed gx and gz.  Returns 0,
 * or -1 on bad input. */
int orc_set_scheme(orc_sim* s, int ko_order, int adv);
/* stencil radius of a discretisation */
int orc_scheme_radius(int ko_order, int adv);
/* Number of flavours (2 or 3); 3 allocates the 10 extra fields (zero) and
 * sets the vacuum Hamiltonian hv.  Call before orc_fill_init.  Returns 0, or
 * -1 on bad input. */
int orc_set_flavors(orc_sim* s, int nflavor);
/* The 3-flavour vacuum Hamiltonian of nuosc_class.cpp:543-569 into hv[9]. */
void orc_vacuum_h3(double hv[9]);

/* nuosc_init.cpp:17-137.  ipt 0,2,3 as the reference; ipt 5 = seeded complex
 * perturbation (counter-based RNG keyed by seed and the interior index).
 * Returns 0 on success, -1 for an unsupported ipt. */
int orc_fill_init(orc_sim* s, int ipt, double alpha, double eps0, double sigma,
                  double lnue, double lnueb, double lnuex, double lnuebx, uint64_t seed);

/* The two uniforms in [-1,1) used by ipt 5 for interior point (i,j,v). */
void orc_seed_uniforms(uint64_t seed, long lin, double* ur, double* ui);

long orc_idx(const orc_sim* s, int i, int j, int v);

void orc_periodic_ghosts(orc_sim* s, double* const f[ORC_NVAR]);
/* Open z boundary at time t, x periodic.  For ghost cell z (centre
 * z0 + (j+1/2)dz, j < 0 or j >= nz) and direction v: if v enters the domain
 * there (vz > 0 on the z0 side, vz < 0 on the z1 side) the ghost holds the
 * ipt-0 state of amplitude a = eps0 exp(-(z - vz t - zc)^2 / (2 sigma^2))
 * (a = 0, the unperturbed state, unless fill_init used ipt 0); otherwise it
 * is the quadratic extrapolation of the three nearest interior cells
 * (coefficients (3,-3,1), (6,-8,3), (10,-15,6) for ghost depth 1, 2, 3). */
void orc_open_ghosts(orc_sim* s, double* const f[ORC_NVAR], double t);
/* The boundary fill of the configured mode at time t. */
void orc_fill_ghosts(orc_sim* s, double* const f[ORC_NVAR], double t);
void orc_calrhs(const orc_sim* s, double* const out[ORC_NVAR], double* const in[ORC_NVAR]);
void orc_step_rk4(orc_sim* s);
/* SSP-RK3 (nuosc_class.cpp:429-452) with a boundary fill before each RHS. */
void orc_step_srk3(orc_sim* s);
void orc_renormalize(orc_sim* s, double* const f[ORC_NVAR]);

/* Diagnostics (nuosc_ana.cpp:34-137).  out[9] = {phy_time, maxdP, surv, survb,
 * avgP, avgPb, aM0, Lex, ELNe}.  Returns 0, or 1 when maxdP >= 10 (the
 * reference's assert) or a sum is not finite. */
int orc_analysis(orc_sim* s, double out[9]);
/* Write the analysis.dat line for out[] with the reference's stream formatting;
 * returns the number of characters written to buf (buf >= 512 bytes). */
int orc_format_analysis_line(orc_sim* s, const double out[9], char* buf, int buflen);

/* Field access in the interior layout [nx][nz][nv] (f in 0..nvar-1, the
 * field order of ORC_NVAR). */
void orc_get_field(const orc_sim* s, int which_buf, int f, double* dst);
void orc_set_field(orc_sim* s, int which_buf, int f, const double* src);
/* Whole ghost-inclusive buffer (lpts doubles). */
double* orc_field_ptr(orc_sim* s, int which_buf, int f);

#ifdef __cplusplus
}
#endif
#endif
