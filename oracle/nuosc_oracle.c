/*
 * nuosc_oracle.c -- CPU restatement of the nuosc RK4 core.  TEST INFRASTRUCTURE
 * (the parity checker); see nuosc_oracle.h for scope, pinning and the rule
 * that only tests, smoke() and bench.py's cpu_baseline use it.
 *
 * Arithmetic follows the reference's expression order term by term (the
 * comments cite the reference lines).  Build with -ffp-contract=off so that no
 * multiply-add is fused, as in the reference's x86-64 build.
 */
#include "nuosc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* 3-vector helpers (Point.h:2-62): cross product, dot, norm                  */
/* ------------------------------------------------------------------------- */
typedef struct { double x, y, z; } p3;

static p3 p3_make(double x, double y, double z) { p3 r; r.x = x; r.y = y; r.z = z; return r; }
static p3 p3_add(p3 a, p3 b) { return p3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static p3 p3_sub(p3 a, p3 b) { return p3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static p3 p3_scale(double k, p3 a) { return p3_make(k * a.x, k * a.y, k * a.z); }
/* Point.h:34-36 (operator* of two points is the cross product) */
static p3 p3_cross(p3 a, p3 b) {
    return p3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static double p3_norm2(p3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static double p3_norm(p3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

/* Point.h:64-69: circumcentre of the planar triangle ABC */
static p3 circumcentre(p3 A, p3 B, p3 C) {
    p3 ABxAC = p3_cross(p3_sub(B, A), p3_sub(C, A));
    p3 AC = p3_sub(C, A);
    p3 AB = p3_sub(B, A);
    p3 d = p3_sub(p3_scale(p3_norm2(AC), AB), p3_scale(p3_norm2(AB), AC));
    return p3_add(A, p3_scale(0.5 / p3_norm2(ABxAC), p3_cross(ABxAC, d)));
}

/* ------------------------------------------------------------------------- */
/* Icosahedral Voronoi velocity grid (icosahedron.h:36-171)                   */
/* ------------------------------------------------------------------------- */
/* Corner table of the 10 rhombic patches (icosahedron.h:39-40): patch p has
 * corners (a,b,c,d) = ICO_PATCH[4p..4p+3] among the 12 base vertices. */
static const unsigned ICO_PATCH[40] = {1, 0, 2, 6,  6, 2, 7, 11, 2, 0, 3, 7,  7, 3, 8, 11,
                                       3, 0, 4, 8,  8, 4, 9, 11, 4, 0, 5, 9,  9, 5, 10, 11,
                                       5, 0, 1, 10, 10, 1, 6, 11};

int orc_vgrid_icosahedron(int nn_in, double* vx, double* vy, double* vz, double* vw) {
    if (nn_in < 1) return -1;
    const unsigned NP = 5, nn = (unsigned)nn_in;
    const double lambda = 1.10714871779409; /* icosahedron.h:41 */
    const unsigned N = 2 * NP * nn * nn + 2;
    p3 bp[12];
    /* icosahedron.h:48-60 */
    for (unsigned i = 1; i <= NP; ++i) {
        bp[i] = p3_make(cos(2 * i * M_PI / NP) * sin(lambda), sin(2 * i * M_PI / NP) * sin(lambda),
                        cos(lambda));
        bp[i + NP] = p3_make(cos((2 * i + 1) * M_PI / NP) * sin(lambda),
                             sin((2 * i + 1) * M_PI / NP) * sin(lambda), -bp[i].z);
    }
    bp[0] = p3_make(0, 0, 1);
    bp[2 * NP + 1] = p3_make(0, 0, -1);

    p3* X = (p3*)calloc(N, sizeof(p3));
    unsigned* ui = (unsigned*)calloc((size_t)2 * NP * (nn + 1) * (nn + 1), sizeof(unsigned));
    double* w = (double*)calloc(N, sizeof(double));
    if (!X || !ui || !w) { free(X); free(ui); free(w); return -1; }

    /* refined, normalised points of each patch (icosahedron.h:62-81) */
    X[0] = bp[0];
    X[N - 1] = bp[2 * NP + 1];
    for (unsigned p = 0; p < 2 * NP; ++p) {
        p3 A = bp[ICO_PATCH[4 * p]], B = bp[ICO_PATCH[4 * p + 1]];
        p3 C = bp[ICO_PATCH[4 * p + 2]], D = bp[ICO_PATCH[4 * p + 3]];
        for (unsigned i = 0; i < nn; ++i)
            for (unsigned j = 0; j < nn; ++j) {
                unsigned pij = p * nn * nn + i * nn + j + 1;
                double di = (double)i, dj = (double)j, dnj = (double)(nn - j), dni = (double)(nn - i);
                p3 q = p3_add(p3_scale(di, p3_add(p3_scale(dj, C), p3_scale(dnj, D))),
                              p3_scale(dni, p3_add(p3_scale(dj, B), p3_scale(dnj, A))));
                double iL = 1.0 / p3_norm(q);
                X[pij] = p3_scale(iL, q);
            }
    }

#define FID(p, i, j) ((((p) * (nn + 1) + (i)) * (nn + 1)) + (j))
    /* unique-point stitching of the patch borders (icosahedron.h:107-135) */
    for (unsigned p = 0; p < 2 * NP; ++p) {
        for (unsigned i = 0; i < nn; ++i)
            for (unsigned j = 0; j < nn; ++j) ui[FID(p, i, j)] = p * nn * nn + i * nn + j + 1;
        if (p == 2 * NP - 1) {
            for (unsigned i = 1; i <= nn; ++i) {
                ui[FID(p, i - 1, nn)] = 1 + nn * (i - 1);
                ui[FID(p, nn, i)] = 2 * nn * nn - (nn - 1) - nn * (i - 1);
            }
            ui[FID(p, nn, 0)] = N - 1;
        } else if (p % 2 == 1) {
            for (unsigned i = 1; i <= nn; ++i) {
                ui[FID(p, i - 1, nn)] = (p + 1) * nn * nn + 1 + nn * (i - 1);
                ui[FID(p, nn, i)] = (p + 3) * nn * nn - (nn - 1) - nn * (i - 1);
            }
            ui[FID(p, nn, 0)] = N - 1;
        } else {
            for (unsigned i = 1; i <= nn; ++i) {
                ui[FID(p, i, nn)] = ((p + 2) * nn * nn + (nn - i) + 1) % (2 * NP * nn * nn);
                ui[FID(p, nn, i - 1)] = (p + 1) * nn * nn + i;
            }
            ui[FID(p, 0, nn)] = 0;
        }
    }

    /* Voronoi weights: each grid triangle gives each corner the planar areas
     * of the two sub-triangles (corner, edge midpoint, circumcentre) that touch
     * it (icosahedron.h:137-170).  Accumulated in the same order. */
    for (unsigned p = 0; p < 2 * NP; ++p)
        for (unsigned i = 0; i < nn; ++i)
            for (unsigned j = 0; j < nn; ++j) {
                unsigned iA = ui[FID(p, i, j)], iC = ui[FID(p, i + 1, j + 1)];
                p3 A = X[iA], C = X[iC];
                for (int t = 0; t < 2; ++t) {
                    unsigned iB = (t == 0) ? ui[FID(p, i, j + 1)] : ui[FID(p, i + 1, j)];
                    p3 B = X[iB];
                    p3 O = circumcentre(A, B, C);
                    double fAB = p3_norm(p3_scale(0.25, p3_cross(p3_sub(O, A), p3_sub(B, A))));
                    double fBC = p3_norm(p3_scale(0.25, p3_cross(p3_sub(O, B), p3_sub(C, B))));
                    double fCA = p3_norm(p3_scale(0.25, p3_cross(p3_sub(O, C), p3_sub(A, C))));
                    w[iA] += fAB + fCA;
                    w[iB] += fBC + fAB;
                    w[iC] += fCA + fBC;
                }
            }
#undef FID
    for (unsigned k = 0; k < N; ++k) {
        vx[k] = X[k].x;
        vy[k] = X[k].y;
        vz[k] = X[k].z;
        vw[k] = w[k];
    }
    free(X);
    free(ui);
    free(w);
    return (int)N;
}

/* ------------------------------------------------------------------------- */
/* Simulation object                                                          */
/* ------------------------------------------------------------------------- */
long orc_idx(const orc_sim* s, int i, int j, int v) {
    /* CartGrid.h:91 */
    return ((long)(i + s->gx) * (s->nz + 2 * s->gz) + j + s->gz) * s->nv + v;
}

static double** bufsel(orc_sim* s, int which) {
    switch (which) {
        case 0: return s->stat;
        case 1: return s->rhs;
        case 2: return s->pre;
        default: return s->cor;
    }
}

orc_sim* orc_create(int nn, int gx, int gz, double x0, double x1, double z0, double z1, double dx,
                    double dz, double cfl, double ko, int vgrid_mode) {
    if (nn < 1 || gx < 2 || gz < 2 || !(dx > 0) || !(dz > 0)) return NULL;
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    if (!s) return NULL;
    s->gx = gx;
    s->gz = gz;
    s->x0 = x0; s->x1 = x1; s->z0 = z0; s->z1 = z1;
    s->dx = dx; s->dz = dz;
    /* CartGrid.cpp:197-201 */
    s->nx = (int)((x1 - x0) / dx);
    s->nz = (int)((z1 - z0) / dz);
    if (s->nx < 1 || s->nz < 1) { free(s); return NULL; }
    s->X = (double*)malloc(sizeof(double) * s->nx);
    s->Z = (double*)malloc(sizeof(double) * s->nz);
    for (int i = 0; i < s->nx; ++i) s->X[i] = x0 + (i + 0.5) * dx;
    for (int i = 0; i < s->nz; ++i) s->Z[i] = z0 + (i + 0.5) * dz;

    int N = 10 * nn * nn + 2;
    s->vx = (double*)malloc(sizeof(double) * N);
    s->vy = (double*)malloc(sizeof(double) * N);
    s->vz = (double*)malloc(sizeof(double) * N);
    s->vw = (double*)malloc(sizeof(double) * N);
    s->nv = orc_vgrid_icosahedron(nn, s->vx, s->vy, s->vz, s->vw);
    if (vgrid_mode == 1)
        for (int k = 0; k < N; ++k) s->vw[k] = 4.0 * M_PI / N;

    s->lpts = (long)(s->nx + 2 * gx) * (s->nz + 2 * gz) * s->nv;
    /* nuosc_class.cpp:461-464 */
    s->ds_L = dx * dz / (z1 - z0) / (x1 - x0);
    s->cfl = cfl;
    s->dt = dz * cfl;
    s->ko = ko;
    s->ko_order = 3;
    s->adv = ORC_ADV_CFD4;
    s->phy_time = 0.0;
    /* nuosc_class.h:120-126 */
    const double theta = 37 * M_PI / 180.;
    s->ct = cos(2 * theta);
    s->st = sin(2 * theta);
    s->pmo = 0.1;
    s->mu = 1.0;
    s->renorm = 0;
    s->G0 = (double*)calloc(s->nv, sizeof(double));
    s->G0b = (double*)calloc(s->nv, sizeof(double));
    s->nvar = ORC_NVAR2;
    for (int f = 0; f < ORC_NVAR2; ++f) {
        s->stat[f] = (double*)calloc(s->lpts, sizeof(double));
        s->rhs[f] = (double*)calloc(s->lpts, sizeof(double));
        s->pre[f] = (double*)calloc(s->lpts, sizeof(double));
        s->cor[f] = (double*)calloc(s->lpts, sizeof(double));
        if (!s->stat[f] || !s->rhs[f] || !s->pre[f] || !s->cor[f]) { orc_destroy(s); return NULL; }
    }
    return s;
}

void orc_destroy(orc_sim* s) {
    if (!s) return;
    for (int f = 0; f < ORC_NVAR; ++f) {
        free(s->stat[f]); free(s->rhs[f]); free(s->pre[f]); free(s->cor[f]);
    }
    free(s->X); free(s->Z); free(s->vx); free(s->vy); free(s->vz); free(s->vw);
    free(s->G0); free(s->G0b);
    free(s);
}

void orc_set_physics(orc_sim* s, double mu, double pmo, int renorm) {
    s->mu = mu;
    s->pmo = pmo;
    s->renorm = renorm;
}

int orc_set_boundary(orc_sim* s, int bc, double zc) {
    if (bc != ORC_BC_PERIODIC && bc != ORC_BC_OPEN_Z) return -1;
    if (bc == ORC_BC_OPEN_Z && (s->nz < 3 || s->gz > 3)) return -1;
    s->bc = bc;
    s->zc = zc;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Initial data (nuosc_init.cpp:4-137)                                        */
/* ------------------------------------------------------------------------- */
/* splitmix64 finaliser (Steele, Lea & Flood 2014), used as a counter-based
 * generator: value = mix(key + counter). */
static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

void orc_seed_uniforms(uint64_t seed, long lin, double* ur, double* ui) {
    uint64_t key = splitmix64(seed ^ 0x6A09E667F3BCC909ull);
    uint64_t h1 = splitmix64(key + 2ull * (uint64_t)lin);
    uint64_t h2 = splitmix64(key + 2ull * (uint64_t)lin + 1ull);
    *ur = 2.0 * ((double)(h1 >> 11) * 0x1.0p-53) - 1.0;
    *ui = 2.0 * ((double)(h2 >> 11) * 0x1.0p-53) - 1.0;
}

/* nuosc_init.cpp:8-10: angular Gaussian centred at (vx0,vz0)=(1,1) */
static double g2(double vx, double vz, double sx, double sz) {
    return exp(-(vx - 1.0) * (vx - 1.0) / (2.0 * sx * sx) - (vz - 1.0) * (vz - 1.0) / (2.0 * sz * sz));
}

int orc_fill_init(orc_sim* s, int ipt, double alpha, double eps0, double sigma, double lnue,
                  double lnueb, double lnuex, double lnuebx, uint64_t seed) {
    if (!(ipt == 0 || ipt == 2 || ipt == 3 || ipt == 5)) return -1;
    const int nv = s->nv;
    double* ng = (double*)malloc(sizeof(double) * nv);
    double* ngb = (double*)malloc(sizeof(double) * nv);
    /* normalisation (nuosc_init.cpp:74-87), sequential order */
    double ing0 = 0, ing1 = 0;
    for (int v = 0; v < nv; ++v) {
        ng[v] = g2(s->vx[v], s->vz[v], lnuex, lnue);
        ngb[v] = g2(s->vx[v], s->vz[v], lnuebx, lnueb);
        ing0 += s->vw[v] * ng[v];
        ing1 += s->vw[v] * ngb[v];
    }
    ing0 = 1.0 / ing0;
    ing1 = 1.0 / ing1;
    for (int v = 0; v < nv; ++v) {
        s->G0[v] = ng[v] * ing0;
        s->G0b[v] = alpha * ngb[v] * ing1;
    }
    free(ng);
    free(ngb);

    double n00 = 0, n01 = 0;
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->nz; ++j)
            for (int v = 0; v < nv; ++v) {
                long ijv = orc_idx(s, i, j, v);
                const double G0 = s->G0[v], G0b = s->G0b[v];
                double tmpr, tmpi = 0.0, z = s->Z[j];
                if (ipt == 0)      /* eps_c, nuosc_init.cpp:4 (centre zc, 0 there) */
                    tmpr = eps0 * exp(-(z - s->zc) * (z - s->zc) / (2.0 * sigma * sigma));
                else if (ipt == 2) /* eps_p, nuosc_init.cpp:6 */
                    tmpr = eps0 * (1.0 + cos(2 * M_PI * (z - 0.0) / (2.0 * sigma * sigma))) * 0.5;
                else if (ipt == 3)
                    tmpr = eps0;
                else {             /* seeded complex perturbation */
                    double ur, uim;
                    orc_seed_uniforms(seed, ((long)i * s->nz + j) * nv + v, &ur, &uim);
                    tmpr = eps0 * ur;
                    tmpi = eps0 * uim;
                }
                double p3o = sqrt(1.0 - tmpr * tmpr - tmpi * tmpi);
                s->stat[0][ijv] = 0.5 * G0 * (1.0 + p3o);
                s->stat[1][ijv] = 0.5 * G0 * (1.0 - p3o);
                s->stat[2][ijv] = 0.5 * G0 * tmpr;
                s->stat[3][ijv] = 0.5 * G0 * tmpi;
                s->stat[4][ijv] = 0.5 * G0b * (1.0 + p3o);
                s->stat[5][ijv] = 0.5 * G0b * (1.0 - p3o);
                s->stat[6][ijv] = 0.5 * G0b * tmpr;
                s->stat[7][ijv] = -0.5 * G0b * tmpi;
                n00 += s->vw[v] * s->stat[0][ijv];
                n01 += s->vw[v] * s->stat[4][ijv];
            }
    /* the reference's ipt 0/2/3 store emi = bemi = +0.0 exactly */
    if (ipt != 5) {
        for (int i = 0; i < s->nx; ++i)
            for (int j = 0; j < s->nz; ++j)
                for (int v = 0; v < nv; ++v) {
                    long ijv = orc_idx(s, i, j, v);
                    s->stat[3][ijv] = 0.0;
                    s->stat[7][ijv] = 0.0;
                }
    }
    /* the reference's fillInitValue writes the 8 two-flavour fields only; the
     * tau fields keep their construction value, zero (nuosc_class.h:24-46) */
    for (int q = ORC_NVAR2; q < s->nvar; ++q)
        for (int i = 0; i < s->nx; ++i)
            for (int j = 0; j < s->nz; ++j)
                memset(s->stat[q] + orc_idx(s, i, j, 0), 0, sizeof(double) * nv);
    s->n_nue0[0] = n00 * s->ds_L;
    s->n_nue0[1] = n01 * s->ds_L;
    s->phy_time = 0.0;
    s->analysis_lines = 0;
    s->inflow_on = (ipt == 0);
    s->inflow_eps0 = eps0;
    s->inflow_sigma = sigma;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Periodic ghost fill (nuosc_boundary.cpp:4-52); corners are left alone.     */
/* ------------------------------------------------------------------------- */
void orc_periodic_ghosts(orc_sim* s, double* const f[ORC_NVAR]) {
    const int nx = s->nx, nz = s->nz, nv = s->nv, gx = s->gx, gz = s->gz;
    /* sequential over j so that ghost-from-ghost copies (nz < gz) resolve in
     * the reference's loop order */
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < gz; ++j)
            for (int v = 0; v < nv; ++v) {
                long l0 = orc_idx(s, i, -j - 1, v), l1 = orc_idx(s, i, nz - j - 1, v);
                long r0 = orc_idx(s, i, nz + j, v), r1 = orc_idx(s, i, j, v);
                for (int q = 0; q < s->nvar; ++q) f[q][l0] = f[q][l1];
                for (int q = 0; q < s->nvar; ++q) f[q][r0] = f[q][r1];
            }
    for (int i = 0; i < gx; ++i)
#pragma omp parallel for
        for (int j = 0; j < nz; ++j)
            for (int v = 0; v < nv; ++v) {
                long l0 = orc_idx(s, -i - 1, j, v), l1 = orc_idx(s, nx - i - 1, j, v);
                long r0 = orc_idx(s, nx + i, j, v), r1 = orc_idx(s, i, j, v);
                for (int q = 0; q < s->nvar; ++q) f[q][l0] = f[q][l1];
                for (int q = 0; q < s->nvar; ++q) f[q][r0] = f[q][r1];
            }
}

/* ------------------------------------------------------------------------- */
/* Open z boundary (no reference counterpart: updateInjetOpenBoundary,        */
/* nuosc_boundary.cpp:54-57, is a stub).                                      */
/* ------------------------------------------------------------------------- */
void orc_open_ghosts(orc_sim* s, double* const f[ORC_NVAR], double t) {
    const int nx = s->nx, nz = s->nz, nv = s->nv, gz = s->gz;
    /* quadratic extrapolation weights of interior cells 0,1,2 (counted inward)
     * for ghost depth 1, 2, 3 */
    static const double W[3][3] = {{3.0, -3.0, 1.0}, {6.0, -8.0, 3.0}, {10.0, -15.0, 6.0}};
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int side = 0; side < 2; ++side)
            for (int q = 1; q <= gz; ++q) {
                const int j = side == 0 ? -q : nz - 1 + q;
                const int e0 = side == 0 ? 0 : nz - 1, dir = side == 0 ? 1 : -1;
                const double z = s->z0 + (j + 0.5) * s->dz;
                for (int v = 0; v < nv; ++v) {
                    const long g = orc_idx(s, i, j, v);
                    const double vz = s->vz[v];
                    const int inflow = side == 0 ? (vz > 0.0) : (vz < 0.0);
                    if (inflow) {
                        double a = 0.0;
                        if (s->inflow_on) {
                            const double zz = z - vz * t - s->zc;
                            a = s->inflow_eps0 * exp(-zz * zz / (2.0 * s->inflow_sigma * s->inflow_sigma));
                        }
                        const double p3o = sqrt(1.0 - a * a);
                        const double G0 = s->G0[v], G0b = s->G0b[v];
                        f[0][g] = 0.5 * G0 * (1.0 + p3o);
                        f[1][g] = 0.5 * G0 * (1.0 - p3o);
                        f[2][g] = 0.5 * G0 * a;
                        f[3][g] = 0.0;
                        f[4][g] = 0.5 * G0b * (1.0 + p3o);
                        f[5][g] = 0.5 * G0b * (1.0 - p3o);
                        f[6][g] = 0.5 * G0b * a;
                        f[7][g] = 0.0;
                        for (int k = ORC_NVAR2; k < s->nvar; ++k) f[k][g] = 0.0;
                    } else {
                        const long c0 = orc_idx(s, i, e0, v), c1 = orc_idx(s, i, e0 + dir, v),
                                   c2 = orc_idx(s, i, e0 + 2 * dir, v);
                        const double* w = W[q - 1];
                        for (int k = 0; k < s->nvar; ++k)
                            f[k][g] = w[0] * f[k][c0] + w[1] * f[k][c1] + w[2] * f[k][c2];
                    }
                }
            }
    /* x ghosts periodic over the interior z rows (nuosc_boundary.cpp:29-50) */
    for (int i = 0; i < s->gx; ++i)
#pragma omp parallel for
        for (int j = 0; j < nz; ++j)
            for (int v = 0; v < nv; ++v) {
                long l0 = orc_idx(s, -i - 1, j, v), l1 = orc_idx(s, nx - i - 1, j, v);
                long r0 = orc_idx(s, nx + i, j, v), r1 = orc_idx(s, i, j, v);
                for (int k = 0; k < s->nvar; ++k) f[k][l0] = f[k][l1];
                for (int k = 0; k < s->nvar; ++k) f[k][r0] = f[k][r1];
            }
}

void orc_fill_ghosts(orc_sim* s, double* const f[ORC_NVAR], double t) {
    if (s->bc == ORC_BC_OPEN_Z)
        orc_open_ghosts(s, f, t);
    else
        orc_periodic_ghosts(s, f);
}

/* ------------------------------------------------------------------------- */
/* Discretisation variants (nuosc_class.cpp:5-8, 112-183, nuosc_class.h:95)  */
/* ------------------------------------------------------------------------- */
int orc_scheme_radius(int ko_order, int adv) {
    return (ko_order == 5 || adv == ORC_ADV_LOPSIDED) ? 3 : 2;
}

int orc_set_scheme(orc_sim* s, int ko_order, int adv) {
    if (ko_order != 3 && ko_order != 5) return -1;
    if (adv < ORC_ADV_CFD4 || adv > ORC_ADV_OFF) return -1;
    const int r = orc_scheme_radius(ko_order, adv);
    if (s->gx < r || s->gz < r) return -1;
    s->ko_order = ko_order;
    s->adv = adv;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Three flavours (N_FLAVOR_3, common.h:10-16)                                */
/* ------------------------------------------------------------------------- */
/* nuosc_class.cpp:543-569, as written there: mass splittings times 1267*2,
 * mixing angles given as degrees/180 and used as radians (cos(33.82/180)),
 * delta_cp = 0.  The names follow the reference: "hmtr" is the (e,tau)
 * element of U diag(0, dm21, dm31) U^T and "hter" the (mu,tau) element. */
void orc_vacuum_h3(double hv[9]) {
    const double d21 = 7.39e-5 * 1267. * 2., d31 = 2.5229e-3 * 1267. * 2.;
    const double a12 = 33.82 / 180., a13 = 8.61 / 180., a23 = 48.3 / 180., dcp = 0.;
    const double c12 = cos(a12), s12 = sin(a12), c13 = cos(a13), s13 = sin(a13), c23 = cos(a23), s23 = sin(a23);
    const double ue2 = s12 * c13;
    const double um2r = c12 * c23 - s12 * s13 * s23 * cos(dcp);
    const double ut2r = -c12 * s23 - s12 * s13 * c23 * cos(dcp);
    const double um2i = -s12 * s13 * s23 * sin(dcp);
    const double ut2i = -s12 * s13 * c23 * sin(dcp);
    const double ue3r = s13 * cos(dcp);
    const double ue3i = -s13 * sin(dcp);
    const double um3 = c13 * s23;
    const double ut3 = c13 * c23;
    double ee = d21 * (ue2 * ue2) + d31 * (s13 * s13);
    double mm = d21 * (um2r * um2r + um2i * um2i) + d31 * (um3 * um3);
    double tt = d21 * (ut2r * ut2r + ut2i * ut2i) + d31 * (ut3 * ut3);
    hv[3] = d21 * (ue2 * um2r) + d31 * (ue3r * um3);
    hv[4] = d21 * (-ue2 * um2i) + d31 * (ue3i * um3);
    hv[5] = d21 * (ue2 * ut2r) + d31 * (ue3r * ut3);
    hv[6] = d21 * (-ue2 * ut2i) + d31 * (ue3i * ut3);
    hv[7] = d21 * (um2r * ut2r + um2i * ut2i) + d31 * (um3 * ut3);
    hv[8] = d21 * (ut2r * um2i - um2r * ut2i);
    const double tr = (ee + mm + tt) / 3.0;
    hv[0] = ee - tr;
    hv[1] = mm - tr;
    hv[2] = tt - tr;
}

int orc_set_flavors(orc_sim* s, int nflavor) {
    if (nflavor == 2) {
        s->nvar = ORC_NVAR2;
        return 0;
    }
    if (nflavor != 3) return -1;
    for (int f = ORC_NVAR2; f < ORC_NVAR3; ++f) {
        if (!s->stat[f]) {
            s->stat[f] = (double*)calloc(s->lpts, sizeof(double));
            s->rhs[f] = (double*)calloc(s->lpts, sizeof(double));
            s->pre[f] = (double*)calloc(s->lpts, sizeof(double));
            s->cor[f] = (double*)calloc(s->lpts, sizeof(double));
        }
        if (!s->stat[f] || !s->rhs[f] || !s->pre[f] || !s->cor[f]) return -1;
    }
    s->nvar = ORC_NVAR3;
    orc_vacuum_h3(s->hv);
    return 0;
}

static int sgn_d(double a) { return (0.0 < a) - (a < 0.0); }

/* superbee limiter (nuosc_class.cpp:5-8) */
static double flimiter(double a, double b) {
    return 0.5 * (sgn_d(a) + sgn_d(b)) * fmax(fmin(2 * fabs(a), fabs(b)), fmin(fabs(a), 2 * fabs(b)));
}

/* Advection term of one field at one point along one axis, x = pointer to the
 * point, st = stride along the axis, v = velocity component, h = spacing,
 * zax = the z axis (nuosc_class.cpp:127-183, non-default branches).  MUSCL is
 * written with one sign that differs between the axes (:150-157): for v < 0
 * the x term ends in + flimiter(x0 - x-1, x1 - x0)/2, the z term in -. */
static double adv_axis(int adv, const double* x, long st, double v, double h, int zax) {
    switch (adv) {
        case ORC_ADV_CFD2: {
            const double f = -v / (2 * h);
            return f * (x[st] - x[-st]);
        }
        case ORC_ADV_UPWIND2: {
            const int sg = sgn_d(v);
            const double f = -sg * v / (2 * h);
            return f * (x[-2 * sg * st] - 4 * x[-sg * st] + 3 * x[0]);
        }
        case ORC_ADV_LOPSIDED: {
            const int sg = sgn_d(v);
            const double f = -sg * v / (12 * h);
            return f * (-x[-3 * sg * st] + 6 * x[-2 * sg * st] - 18 * x[-sg * st] + 10 * x[0] + 3 * x[sg * st]);
        }
        case ORC_ADV_MUSCL: {
            const double up = (v > 0) * (x[0] + 0.5 * flimiter(x[0] - x[-st], x[st] - x[0]) - x[-st] -
                                         0.5 * flimiter(x[-st] - x[-2 * st], x[0] - x[-st]));
            const double lc = 0.5 * flimiter(x[0] - x[-st], x[st] - x[0]);
            const double dn = (v < 0) * (x[st] - 0.5 * flimiter(x[st] - x[0], x[2 * st] - x[st]) - x[0] +
                                         (zax ? -lc : lc));
            return -v / h * (up + dn);
        }
        case ORC_ADV_OFF:
            return 0.0;
        default: {  /* 4th-order central */
            const double f = -v / (12 * h);
            return f * ((x[-2 * st] - x[2 * st]) - 8.0 * (x[-st] - x[st]));
        }
    }
}

/* Kreiss-Oliger term of order 5 along one axis, eps = -ko/h/64
 * (nuosc_class.cpp:112-117) */
static double ko5_axis(const double* x, long st, double eps) {
    return eps * (x[-3 * st] + x[3 * st] - 6 * (x[-2 * st] + x[2 * st]) + 15 * (x[-st] + x[st]) - 20 * x[0]);
}

/* ------------------------------------------------------------------------- */
/* Right-hand side (nuosc_class.cpp:11-307)                                   */
/* ------------------------------------------------------------------------- */
/* Stencil terms of the point x (a pointer into a ghosted field), with the
 * locals of the functions below: advection (nuosc_class.cpp:127-183) and
 * Kreiss-Oliger (:112-123) along z (stride snv) and x (stride nzv), and the
 * projection of a moment vector onto direction v (INP, :187). */
#define ADV_FD(x) (dflt ? (factor_z * (((x)[-2 * snv] - (x)[2 * snv]) - 8.0 * ((x)[-snv] - (x)[snv])) + \
                          factor_x * (((x)[-2 * nzv] - (x)[2 * nzv]) - 8.0 * ((x)[-nzv] - (x)[nzv]))) \
                       : (adv_axis(s->adv, (x), snv, vz[v], s->dz, 1) + adv_axis(s->adv, (x), nzv, vx[v], s->dx, 0)))
#define KO_FD(x) (s->ko_order != 5 \
                      ? (ko_eps_z * ((x)[-2 * snv] + (x)[2 * snv] - 4 * ((x)[-snv] + (x)[snv]) + 6 * (x)[0]) + \
                         ko_eps_x * ((x)[-2 * nzv] + (x)[2 * nzv] - 4 * ((x)[-nzv] + (x)[nzv]) + 6 * (x)[0])) \
                      : (ko5_axis((x), snv, ko5_eps_z) + ko5_axis((x), nzv, ko5_eps_x)))
#define INP(a) ((a)[0] - (a)[1] * vx[v] - (a)[2] * vy[v] - (a)[3] * vz[v])

/* Three flavours: moments nuosc_class.cpp:28-82, interaction :206-252 (the
 * inline-projection branch; the other branch computes the same products),
 * vacuum :260-277, out = I + ADV + KO + vacuum.  Written as the reference
 * has them, including the nubar tau-tau row, whose vacuum term reads the nu
 * fields (:269). */
static void calrhs3(const orc_sim* s, double* const out[ORC_NVAR], double* const in[ORC_NVAR]) {
    const int nx = s->nx, nz = s->nz, nv = s->nv;
    const long snv = nv;
    const long nzv = (long)nv * (nz + 2 * s->gz);
    const double* vx = s->vx; const double* vy = s->vy;
    const double* vz = s->vz; const double* vw = s->vw;
    const double mu = s->mu, pmo = s->pmo;
    const double hee = s->hv[0], hmm = s->hv[1], htt = s->hv[2], hemr = s->hv[3], hemi = s->hv[4];
    const double hmtr = s->hv[5], hmti = s->hv[6], hter = s->hv[7], htei = s->hv[8];
    const double ko_eps_x = -s->ko / s->dx / 16.0;
    const double ko_eps_z = -s->ko / s->dz / 16.0;
    const double ko5_eps_x = -s->ko / s->dx / 64.0;
    const double ko5_eps_z = -s->ko / s->dz / 64.0;
    const int dflt = s->ko_order != 5 && s->adv == ORC_ADV_CFD4;
    enum { EE, MM, EMR, EMI, BEE, BMM, BEMR, BEMI, TT, MTR, MTI, TER, TEI, BTT, BMTR, BMTI, BTER, BTEI };

#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            /* 8 moment vectors: emRm emIp eemm mtRm mtIp teRm teIp mmtt */
            double m[8][4];
            memset(m, 0, sizeof m);
            for (int k = 0; k < nv; ++k) {
                const long p = orc_idx(s, i, j, k);
                const double v4[4] = {1, vx[k], vy[k], vz[k]};
                const double d[8] = {
                    in[BEMR][p] - in[EMR][p],
                    in[BEMI][p] + in[EMI][p],
                    in[BEE][p] - in[BMM][p] - in[EE][p] + in[MM][p],
                    in[BMTR][p] - in[MTR][p],
                    in[BMTI][p] + in[MTI][p],
                    in[BTER][p] - in[TER][p],
                    in[BTEI][p] + in[TEI][p],
                    in[BMM][p] - in[BTT][p] - in[MM][p] + in[TT][p]};
                for (int q = 0; q < 4; ++q)
                    for (int c = 0; c < 8; ++c) m[c][q] += v4[q] * vw[k] * d[c];
            }
            for (int v = 0; v < nv; ++v) {
                const long p = orc_idx(s, i, j, v);
                const double factor_x = -vx[v] / (12 * s->dx);
                const double factor_z = -vz[v] / (12 * s->dz);
                /* projections of the moment vectors onto v */
                const double pA = INP(m[0]), pB = INP(m[1]), pC = INP(m[2]), pD = INP(m[3]);
                const double pE = INP(m[4]), pF = INP(m[5]), pG = INP(m[6]), pH = INP(m[7]);
                const double pTE = -(pC + pH);
                double I[ORC_NVAR3], V[ORC_NVAR3];
                /* nu (fl = 0) and nubar (fl = 1) rows; both see the same moments */
                for (int fl = 0; fl < 2; ++fl) {
                    const int oE = fl ? BEE : EE, oM = fl ? BMM : MM, oT = fl ? BTT : TT;
                    const int oEMR = fl ? BEMR : EMR, oEMI = fl ? BEMI : EMI, oMTR = fl ? BMTR : MTR;
                    const int oMTI = fl ? BMTI : MTI, oTER = fl ? BTER : TER, oTEI = fl ? BTEI : TEI;
                    const double ee = in[oE][p], mm = in[oM][p], tt = in[oT][p];
                    const double emr = in[oEMR][p], emi = in[oEMI][p], mtr = in[oMTR][p];
                    const double mti = in[oMTI][p], ter = in[oTER][p], tei = in[oTEI][p];
                    I[oE] = 2 * mu * (emi * pA + emr * pB - tei * pF - ter * pG);
                    I[oM] = 2 * mu * (-emi * pA - emr * pB + mti * pD + mtr * pE);
                    I[oT] = 2 * mu * (-mti * pD - mtr * pE + tei * pF + ter * pG);
                    I[oEMR] = -mu * ((ee - mm) * pB + pC * emi - pE * ter + pD * tei - mti * pF + mtr * pG);
                    I[oEMI] = mu * ((mm - ee) * pA + pC * emr - pE * tei - pD * ter + mti * pG + mtr * pF);
                    I[oMTR] = -mu * ((mm - tt) * pE + pB * ter - pA * tei + emi * pF - emr * pG + mti * pH);
                    I[oMTI] = mu * ((tt - mm) * pD + pB * tei + pA * ter - emi * pG - emr * pF + mtr * pH);
                    I[oTER] = -mu * ((tt - ee) * pG - pB * mtr + pA * mti - emi * pD + emr * pE + tei * pTE);
                    I[oTEI] = mu * ((ee - tt) * pF - pB * mti - pA * mtr + emi * pE + emr * pD + ter * pTE);
                    V[oE] = -pmo * 2 * (emi * hemr + emr * hemi - htei * ter + hter * tei);
                    V[oM] = pmo * 2 * (emi * hemr - emr * hemi + hmti * mtr - hmtr * mti);
                    V[oEMR] = pmo * ((mm - ee) * hemi + emi * (hee - hmm) + hmti * ter + hmtr * tei - htei * mtr - hter * mti);
                    V[oEMI] = pmo * ((ee - mm) * hemr - emr * (hee - hmm) - hmti * tei + hmtr * ter + htei * mti - hter * mtr);
                    V[oTER] = pmo * ((ee - tt) * htei - emr * hmti - emi * hmtr + hemi * mtr + hemr * mti - tei * (hee - htt));
                    V[oTEI] = pmo * ((tt - ee) * hter - emr * hmtr + emi * hmti - hemi * mti + hemr * mtr + ter * (hee - htt));
                    if (fl == 0) {
                        V[oMTR] = pmo * ((tt - mm) * hmti + emr * htei + emi * hter - hemi * ter - hemr * tei + mti * (hmm - htt));
                        V[oMTI] = pmo * ((mm - tt) * hmtr + emr * hter - emi * htei + hemi * tei - hemr * ter - mtr * (hmm - htt));
                    } else {   /* the same terms in the order of :270-271 */
                        V[oMTR] = pmo * ((tt - mm) * hmti + emr * htei + emi * hter - hemr * tei - hemi * ter + mti * (hmm - htt));
                        V[oMTI] = pmo * ((mm - tt) * hmtr + emr * hter - emi * htei - hemr * ter + hemi * tei - mtr * (hmm - htt));
                    }
                }
                {   /* tau-tau rows (:268-269): both read the nu fields */
                    const double mtr = in[MTR][p], mti = in[MTI][p], ter = in[TER][p], tei = in[TEI][p];
                    V[TT] = -pmo * 2 * (hmti * mtr + hmtr * mti + htei * ter - hter * tei);
                    V[BTT] = V[TT];
                }
                for (int f = 0; f < ORC_NVAR3; ++f) {
                    const double* x = &in[f][p];
                    /* the vacuum term enters as written: "- pmo*2*(...)" rows subtract */
                    out[f][p] = I[f] + ADV_FD(x) + KO_FD(x) + V[f];
                }
            }
        }
}

void orc_calrhs(const orc_sim* s, double* const out[ORC_NVAR], double* const in[ORC_NVAR]) {
    if (s->nvar == ORC_NVAR3) {
        calrhs3(s, out, in);
        return;
    }
    const int nx = s->nx, nz = s->nz, nv = s->nv;
    const long snv = nv;                                  /* z stride */
    const long nzv = (long)nv * (nz + 2 * s->gz);         /* x stride */
    const double* vx = s->vx; const double* vy = s->vy;
    const double* vz = s->vz; const double* vw = s->vw;
    const double mu = s->mu, pmo = s->pmo, ct = s->ct, st = s->st;
    /* nuosc_class.cpp:120-121 (order 3), 114-115 (order 5) */
    const double ko_eps_x = -s->ko / s->dx / 16.0;
    const double ko_eps_z = -s->ko / s->dz / 16.0;
    const double ko5_eps_x = -s->ko / s->dx / 64.0;
    const double ko5_eps_z = -s->ko / s->dz / 64.0;
    const int dflt = s->ko_order != 5 && s->adv == ORC_ADV_CFD4;

#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            /* velocity moments (nuosc_class.cpp:28-82), sequential in k */
            double emRm[4] = {0, 0, 0, 0}, emIp[4] = {0, 0, 0, 0}, eemm[4] = {0, 0, 0, 0};
            for (int k = 0; k < nv; ++k) {
                long ijk = orc_idx(s, i, j, k);
                const double v4[4] = {1, vx[k], vy[k], vz[k]};
                for (int q = 0; q < 4; ++q) {
                    emRm[q] += v4[q] * vw[k] * (in[6][ijk] - in[2][ijk]);
                    emIp[q] += v4[q] * vw[k] * (in[7][ijk] + in[3][ijk]);
                    eemm[q] += v4[q] * vw[k] * (in[4][ijk] - in[5][ijk] - in[0][ijk] + in[1][ijk]);
                }
            }
            for (int v = 0; v < nv; ++v) {
                const long ijv = orc_idx(s, i, j, v);
                /* nuosc_class.cpp:180-183 */
                const double factor_x = -vx[v] / (12 * s->dx);
                const double factor_z = -vz[v] / (12 * s->dz);
                const double* ee = &in[0][ijv];
                const double* mm = &in[1][ijv];
                const double* emr = &in[2][ijv];
                const double* emi = &in[3][ijv];
                const double* bee = &in[4][ijv];
                const double* bmm = &in[5][ijv];
                const double* bemr = &in[6][ijv];
                const double* bemi = &in[7][ijv];
                /* interaction (nuosc_class.cpp:190-200) */
                const double iemRm = INP(emRm);
                const double iemIp = INP(emIp);
                const double ieemm = INP(eemm);
                const double Iee = 2 * mu * (emi[0] * iemRm + emr[0] * iemIp);
                const double Imm = -Iee;
                const double Iemr = -mu * ((ee[0] - mm[0]) * iemIp + ieemm * emi[0]);
                const double Iemi = mu * ((mm[0] - ee[0]) * iemRm + ieemm * emr[0]);
                const double Ibee = 2 * mu * (bemi[0] * iemRm + bemr[0] * iemIp);
                const double Ibemr = -mu * ((bee[0] - bmm[0]) * iemIp + INP(eemm) * bemi[0]);
                const double Ibemi = mu * ((bmm[0] - bee[0]) * iemRm + INP(eemm) * bemr[0]);
                /* nuosc_class.cpp:280-287 (the bar mm row uses -Ibee) */
                out[0][ijv] = Iee + ADV_FD(ee) + KO_FD(ee) - pmo * 2 * st * emi[0];
                out[1][ijv] = Imm + ADV_FD(mm) + KO_FD(mm) + pmo * 2 * st * emi[0];
                out[2][ijv] = Iemr + ADV_FD(emr) + KO_FD(emr) - pmo * 2 * ct * emi[0];
                out[3][ijv] = Iemi + ADV_FD(emi) + KO_FD(emi) + pmo * (2 * ct * emr[0] + st * (ee[0] - mm[0]));
                out[4][ijv] = Ibee + ADV_FD(bee) + KO_FD(bee) - pmo * 2 * st * bemi[0];
                out[5][ijv] = -Ibee + ADV_FD(bmm) + KO_FD(bmm) + pmo * 2 * st * bemi[0];
                out[6][ijv] = Ibemr + ADV_FD(bemr) + KO_FD(bemr) - pmo * 2 * ct * bemi[0];
                out[7][ijv] = Ibemi + ADV_FD(bemi) + KO_FD(bemi) + pmo * (2 * ct * bemr[0] + st * (bee[0] - bmm[0]));
            }
        }
}
#undef ADV_FD
#undef KO_FD
#undef INP

/* ------------------------------------------------------------------------- */
/* RK4 (nuosc_class.cpp:312-351, 374-427)                                     */
/* ------------------------------------------------------------------------- */
/* v0 = v1 + a*v2 over the interior */
static void vec1(orc_sim* s, double** v0, double** v1, double a, double** v2) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->nz; ++j)
            for (int v = 0; v < s->nv; ++v) {
                long k = orc_idx(s, i, j, v);
                for (int f = 0; f < s->nvar; ++f) v0[f][k] = v1[f][k] + a * v2[f][k];
            }
}
/* v0 = v1 + a*(v2 + v3) over the interior */
static void vec2(orc_sim* s, double** v0, double** v1, double a, double** v2, double** v3) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->nz; ++j)
            for (int v = 0; v < s->nv; ++v) {
                long k = orc_idx(s, i, j, v);
                for (int f = 0; f < s->nvar; ++f) v0[f][k] = v1[f][k] + a * (v2[f][k] + v3[f][k]);
            }
}
/* v0 = a*v1 + b*(v2 + dt*v3) over the interior (nuosc_class.cpp:353-372) */
static void vec3(orc_sim* s, double** v0, double a, double** v1, double b, double** v2, double dt, double** v3) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->nz; ++j)
            for (int v = 0; v < s->nv; ++v) {
                long k = orc_idx(s, i, j, v);
                for (int f = 0; f < s->nvar; ++f) v0[f][k] = a * v1[f][k] + b * (v2[f][k] + dt * v3[f][k]);
            }
}
static void swapbuf(double** a, double** b) {
    for (int f = 0; f < ORC_NVAR; ++f) { double* t = a[f]; a[f] = b[f]; b[f] = t; }
}

void orc_step_rk4(orc_sim* s) {
    const double dt = s->dt, t = s->phy_time;
    /* stage 1 (boundary data at the stage times t, t+dt/2, t+dt/2, t+dt) */
    orc_fill_ghosts(s, s->stat, t);
    orc_calrhs(s, s->rhs, s->stat);
    vec1(s, s->pre, s->stat, 0.5 * dt, s->rhs);
    /* stage 2 */
    orc_fill_ghosts(s, s->pre, t + 0.5 * dt);
    orc_calrhs(s, s->cor, s->pre);
    vec1(s, s->rhs, s->rhs, 2.0, s->cor);
    vec1(s, s->cor, s->stat, 0.5 * dt, s->cor);
    swapbuf(s->pre, s->cor);
    /* stage 3 */
    orc_fill_ghosts(s, s->pre, t + 0.5 * dt);
    orc_calrhs(s, s->cor, s->pre);
    vec1(s, s->rhs, s->rhs, 2.0, s->cor);
    vec1(s, s->cor, s->stat, dt, s->cor);
    swapbuf(s->pre, s->cor);
    /* stage 4 */
    orc_fill_ghosts(s, s->pre, t + dt);
    orc_calrhs(s, s->cor, s->pre);
    vec2(s, s->pre, s->stat, 1.0 / 6.0 * dt, s->cor, s->rhs);
    swapbuf(s->pre, s->stat);
    if (s->renorm) orc_renormalize(s, s->stat);
    s->phy_time += dt;
}

/* SSP-RK3 (nuosc_class.cpp:429-452): same stage combinations and buffer
 * use.  The reference's step_srk3 refreshes no ghost cells before its RHS
 * evaluations; here each stage input gets the boundary fill of step_rk4 (at
 * the stage times t, t+dt, t+dt/2). */
void orc_step_srk3(orc_sim* s) {
    const double dt = s->dt, t = s->phy_time;
    orc_fill_ghosts(s, s->stat, t);
    orc_calrhs(s, s->rhs, s->stat);
    vec1(s, s->pre, s->stat, dt, s->rhs);                              /* u1 = u0 + dt L(u0) */
    orc_fill_ghosts(s, s->pre, t + dt);
    orc_calrhs(s, s->rhs, s->pre);
    vec3(s, s->pre, 3.0 / 4.0, s->stat, 1.0 / 4.0, s->pre, dt, s->rhs); /* u2 */
    orc_fill_ghosts(s, s->pre, t + 0.5 * dt);
    orc_calrhs(s, s->cor, s->pre);
    vec3(s, s->stat, 1.0 / 3.0, s->stat, 2.0 / 3.0, s->pre, dt, s->cor);
    if (s->renorm) orc_renormalize(s, s->stat);
    s->phy_time += dt;
}

/* nuosc_ana.cpp:140-166 (as written there: P1 without the factor 2) */
void orc_renormalize(orc_sim* s, double* const f[ORC_NVAR]) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->nz; ++j)
            for (int v = 0; v < s->nv; ++v) {
                long k = orc_idx(s, i, j, v);
                double iG = 1.0 / s->G0[v], iGb = 1.0 / s->G0b[v];
                double P1 = f[2][k] * iG, P2 = -f[3][k] * iG, P3 = (f[0][k] - f[1][k]) * iG;
                double P1b = f[6][k] * iGb, P2b = f[7][k] * iGb, P3b = (f[4][k] - f[5][k]) * iGb;
                double iP = 1.0 / sqrt(P1 * P1 + P2 * P2 + P3 * P3);
                double iPb = 1.0 / sqrt(P1b * P1b + P2b * P2b + P3b * P3b);
                double tmp = iP * (P3) * s->G0[v];
                double tmpb = iPb * (P3b) * s->G0b[v];
                f[0][k] = s->G0[v] + tmp;
                f[1][k] = s->G0[v] - tmp;
                f[2][k] *= iP;
                f[3][k] *= iP;
                f[4][k] = s->G0b[v] + tmpb;
                f[5][k] = s->G0b[v] - tmpb;
                f[6][k] *= iPb;
                f[7][k] *= iPb;
            }
}

/* ------------------------------------------------------------------------- */
/* Diagnostics (nuosc_ana.cpp:3-137), sequential reduction order              */
/* ------------------------------------------------------------------------- */
int orc_analysis(orc_sim* s, double out[9]) {
    double maxdP = 0.0, avgP = 0.0, avgPb = 0.0, surv = 0.0, survb = 0.0;
    double nor = 0.0, norb = 0.0, aM01 = 0.0, aM02 = 0.0, aM03 = 0.0;
    double** f = s->stat;
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->nz; ++j)
            for (int v = 0; v < s->nv; ++v) {
                long k = orc_idx(s, i, j, v);
                const double G0 = s->G0[v], G0b = s->G0b[v], w = s->vw[v];
                /* eval_conserved, nuosc_ana.cpp:10-27 */
                double iG = 1.0 / G0, iGb = 1.0 / G0b;
                double P1 = 2.0 * f[2][k] * iG;
                double P2 = -2.0 * f[3][k] * iG;
                double P3 = (f[0][k] - f[1][k]) * iG;
                double P1b = 2.0 * f[6][k] * iGb;
                double P2b = 2.0 * f[7][k] * iGb; /* sign as in nuosc_ana.cpp:16 */
                double P3b = (f[4][k] - f[5][k]) * iGb;
                double dP = fabs(1.0 - sqrt(P1 * P1 + P2 * P2 + P3 * P3));
                double dPb = fabs(1.0 - sqrt(P1b * P1b + P2b * P2b + P3b * P3b));
                /* reduction body, nuosc_ana.cpp:64-84 (std::max keeps the first
                 * argument when the comparison is false, so NaN is dropped) */
                double m = (dP < dPb) ? dPb : dP;
                maxdP = (maxdP < m) ? m : maxdP;
                surv += w * f[0][k];
                survb += w * f[4][k];
                avgP += w * G0 * dP;
                avgPb += w * G0b * dPb;
                aM01 += w * (f[2][k] - f[6][k]);
                aM02 += w * (-f[3][k] - f[7][k]);
                aM03 += w * 0.5 * (f[0][k] - f[1][k] - f[4][k] + f[5][k]);
                nor += w * G0;
                norb += w * G0b;
            }
    surv /= nor;
    survb /= norb;
    avgP /= nor;
    avgPb /= norb;
    double aM0 = sqrt(aM01 * aM01 + aM02 * aM02 + aM03 * aM03) * s->ds_L;
    double ELNe = fabs(s->n_nue0[0] * (1.0 - surv) - s->n_nue0[1] * (1.0 - survb)) /
                  (s->n_nue0[0] + s->n_nue0[1]);
    double Lex = aM03 * s->ds_L;
    out[0] = s->phy_time;
    out[1] = maxdP;
    out[2] = surv;
    out[3] = survb;
    out[4] = avgP;
    out[5] = avgPb;
    out[6] = aM0;
    out[7] = Lex;
    out[8] = ELNe;
    int bad = !(maxdP < 10);
    for (int q = 2; q < 9; ++q) bad |= !isfinite(out[q]);
    return bad ? 1 : 0;
}

/* nuosc_ana.cpp:126-129: phy_time goes out before setprecision(13) on the
 * stream's first line (default precision 6); every later value, including
 * later phy_time values, goes out with precision 13; the line ends " \n". */
int orc_format_analysis_line(orc_sim* s, const double out[9], char* buf, int buflen) {
    int prec_t = (s->analysis_lines == 0) ? 6 : 13;
    int n = snprintf(buf, buflen, "%.*g %.13g %.13g %.13g %.13g %.13g %.13g %.13g %.13g \n", prec_t,
                     out[0], out[1], out[2], out[3], out[4], out[5], out[6], out[7], out[8]);
    s->analysis_lines++;
    return n;
}

/* ------------------------------------------------------------------------- */
/* Field access                                                               */
/* ------------------------------------------------------------------------- */
void orc_get_field(const orc_sim* s, int which, int f, double* dst) {
    double** b = bufsel((orc_sim*)s, which);
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->nz; ++j)
            memcpy(dst + ((long)i * s->nz + j) * s->nv, b[f] + orc_idx(s, i, j, 0),
                   sizeof(double) * s->nv);
}

void orc_set_field(orc_sim* s, int which, int f, const double* src) {
    double** b = bufsel(s, which);
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->nz; ++j)
            memcpy(b[f] + orc_idx(s, i, j, 0), src + ((long)i * s->nz + j) * s->nv,
                   sizeof(double) * s->nv);
}

double* orc_field_ptr(orc_sim* s, int which, int f) { return bufsel(s, which)[f]; }
