// nuosc_flavor3.hip -- CDNA4 (gfx950) kernels of the 3-flavour time stepper
// (the reference's compile-time N_FLAVOR_3: 18 fields, nuosc_class.cpp:28-82,
// 189-277; vacuum Hamiltonian :543-569).
//
// Kernels over the state copies (the layout of nuosc_dev.h with nf = 18):
//   k_stage3z   grids with nx > 1: the stage as a z march, 8 wavefronts over
//               4 x rows of one 64-direction chunk, the 18 fields split
//               between two groups of waves along the moment vectors (below);
//   k_stage3    one-row grids (nx = 1; written for 4 x rows per block,
//               launched in its one-row form): one wavefront per (site, 64
//               directions), four direction chunks of the row per block:
//               the site's 18 centre values and 8 moment projections
//               give the interaction and vacuum terms; each field's x
//               neighbours come from the block's LDS (the waves' own rows and
//               one edge row loaded by each wave), its z neighbours straight
//               from HBM/L2; the combination of the RK4 or SSP-RK3 stage is
//               stored, and the 32 moments of the stored values are summed
//               over the wave's directions into a per-(site, chunk) partial
//               record;
//   k_moment_totals3  adds a site's partial records in chunk order: the
//               moments the next stage reads (no extra pass over the state);
//   k_moments3  one wavefront per site: the moments of a whole state (after
//               init or upload, and for calRHS).
// All sums run in a fixed order, so a run's result does not depend on the
// launch.
// k_stage3's logical block order walks x quads inside z tiles of 16 sites, so the
// z neighbours of a block were read a block or two earlier and its x
// neighbours (the next quads) 16 blocks earlier, on the same XCD, and most
// stencil re-reads hit L2 (HBM reads 1.3-1.6x one pass of the stage input at
// 512 x 512 x 642).  The cost is the number of neighbour loads per field
// and lane: with 18 fields a register z-window like the 2-flavour kernel's
// would take 180 VGPRs in FP64.
#include <hip/hip_runtime.h>

#include "nuosc_dev.h"
#include "nuosc_lanes.h"

#include <algorithm>

namespace nuosc {
namespace {

constexpr int kF3Rows = 4;    // x rows (wavefronts) per stage block
constexpr int kF3Tile = 16;   // z sites per tile of the logical block order

enum F3Field { EE, MM, EMR, EMI, BEE, BMM, BEMR, BEMI, TT, MTR, MTI, TER, TEI, BTT, BMTR, BMTI, BTER, BTEI };

typedef const __attribute__((address_space(4))) double cdouble3;

// nuosc_class.cpp:57-71: moment differences of one point
template <typename T>
__device__ __forceinline__ void moment_diffs3(const T (&y)[kNVar3], double (&d)[8]) {
    d[0] = (double)y[BEMR] - (double)y[EMR];
    d[1] = (double)y[BEMI] + (double)y[EMI];
    d[2] = (((double)y[BEE] - (double)y[BMM]) - (double)y[EE]) + (double)y[MM];
    d[3] = (double)y[BMTR] - (double)y[MTR];
    d[4] = (double)y[BMTI] + (double)y[MTI];
    d[5] = (double)y[BTER] - (double)y[TER];
    d[6] = (double)y[BTEI] + (double)y[TEI];
    d[7] = (((double)y[BMM] - (double)y[BTT]) - (double)y[MM]) + (double)y[TT];
}

template <typename T>
__global__ __launch_bounds__(256) void k_moments3(const Geometry g, const T* __restrict__ y,
                                                  double* __restrict__ mom) {
    const int lane = threadIdx.x & 63;
    const long long site = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (site >= (long long)g.nx * g.nz) return;   // whole wavefronts
    double acc[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) acc[q] = 0.0;
    const T* p = y + site_off(g, site);
    for (int v = lane; v < g.nv; v += 64) {
        T x[kNVar3];
#pragma unroll
        for (int f = 0; f < kNVar3; ++f) x[f] = p[(long long)f * g.nvp + v];
        const double w = g.vw[v];
        const double q4[4] = {w, g.vx[v] * w, g.vy[v] * w, g.vz[v] * w};
        double d[8];
        moment_diffs3<T>(x, d);
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int s = 0; s < 4; ++s) acc[4 * c + s] += q4[s] * d[c];
    }
    const double tot = wave_sum32(acc, lane);
    if ((lane & 1) == 0) mom[site * kNMom3 + (lane >> 1)] = tot;
}

template <typename T>
struct F3Args {
    Geometry g;
    const T* yin;
    const T* u;
    T* w;
    T* yout;
    const double* mom;   // moment totals of yin: [site][32]
    double* part;        // partial moments of yout: [site][nch][32] (stages 1-7)
    double cw, cy, ca, cb;
    double mu, pmo;
    double h[9];
    int tiles, quads, nch;
    int one_row;   // nx = 1: the block's waves take kF3Rows direction chunks of that row
    int cgroups;   // chunk groups of the block grid (nch, or ceil(nch / kF3Rows) for one row)
};

// Moment differences (nuosc_class.cpp:57-71) a field's value enters, with
// its sign: {d index, sign, second d index or -1, its sign}
__device__ constexpr int kDiffOf[kNVar3][4] = {
    {2, -1, -1, 0}, {2, +1, 7, -1}, {0, -1, -1, 0}, {1, +1, -1, 0}, {2, +1, -1, 0}, {2, -1, 7, +1},
    {0, +1, -1, 0}, {1, +1, -1, 0}, {7, +1, -1, 0}, {3, -1, -1, 0}, {4, +1, -1, 0}, {5, -1, -1, 0},
    {6, +1, -1, 0}, {7, -1, -1, 0}, {3, +1, -1, 0}, {4, +1, -1, 0}, {5, +1, -1, 0}, {6, +1, -1, 0}};

// Interaction (nuosc_class.cpp:206-252) and vacuum (:260-277) terms of one
// flavour block (nu: o = 0, nubar: o = 1) from the centre values c and the
// moment projections P = {emRm, emIp, eemm, mtRm, mtIp, teRm, teIp, mmtt}.
template <typename T>
__device__ __forceinline__ void local_terms3(const T (&c)[kNVar3], const T (&P)[8], T mu, T pmo, const T (&h)[9],
                                             T (&I)[kNVar3], T (&V)[kNVar3]) {
    const T hee = h[0], hmm = h[1], htt = h[2], hemr = h[3], hemi = h[4], hmtr = h[5], hmti = h[6], hter = h[7],
            htei = h[8];
    const T pA = P[0], pB = P[1], pC = P[2], pD = P[3], pE = P[4], pF = P[5], pG = P[6], pH = P[7];
    const T pTE = -(pC + pH);
    const T mu2 = (T)2 * mu, pmo2 = pmo * (T)2;
#pragma unroll
    for (int fl = 0; fl < 2; ++fl) {
        const int oE = fl ? BEE : EE, oM = fl ? BMM : MM, oT = fl ? BTT : TT, oEMR = fl ? BEMR : EMR;
        const int oEMI = fl ? BEMI : EMI, oMTR = fl ? BMTR : MTR, oMTI = fl ? BMTI : MTI, oTER = fl ? BTER : TER;
        const int oTEI = fl ? BTEI : TEI;
        const T ee = c[oE], mm = c[oM], tt = c[oT], emr = c[oEMR], emi = c[oEMI], mtr = c[oMTR], mti = c[oMTI],
                ter = c[oTER], tei = c[oTEI];
        I[oE] = mu2 * (emi * pA + emr * pB - tei * pF - ter * pG);
        I[oM] = mu2 * (-emi * pA - emr * pB + mti * pD + mtr * pE);
        I[oT] = mu2 * (-mti * pD - mtr * pE + tei * pF + ter * pG);
        I[oEMR] = -mu * ((ee - mm) * pB + pC * emi - pE * ter + pD * tei - mti * pF + mtr * pG);
        I[oEMI] = mu * ((mm - ee) * pA + pC * emr - pE * tei - pD * ter + mti * pG + mtr * pF);
        I[oMTR] = -mu * ((mm - tt) * pE + pB * ter - pA * tei + emi * pF - emr * pG + mti * pH);
        I[oMTI] = mu * ((tt - mm) * pD + pB * tei + pA * ter - emi * pG - emr * pF + mtr * pH);
        I[oTER] = -mu * ((tt - ee) * pG - pB * mtr + pA * mti - emi * pD + emr * pE + tei * pTE);
        I[oTEI] = mu * ((ee - tt) * pF - pB * mti - pA * mtr + emi * pE + emr * pD + ter * pTE);
        V[oE] = -(pmo2 * (emi * hemr + emr * hemi - htei * ter + hter * tei));
        V[oM] = pmo2 * (emi * hemr - emr * hemi + hmti * mtr - hmtr * mti);
        V[oEMR] = pmo * ((mm - ee) * hemi + emi * (hee - hmm) + hmti * ter + hmtr * tei - htei * mtr - hter * mti);
        V[oEMI] = pmo * ((ee - mm) * hemr - emr * (hee - hmm) - hmti * tei + hmtr * ter + htei * mti - hter * mtr);
        V[oTER] = pmo * ((ee - tt) * htei - emr * hmti - emi * hmtr + hemi * mtr + hemr * mti - tei * (hee - htt));
        V[oTEI] = pmo * ((tt - ee) * hter - emr * hmtr + emi * hmti - hemi * mti + hemr * mtr + ter * (hee - htt));
        if (fl == 0) {
            V[oMTR] = pmo * ((tt - mm) * hmti + emr * htei + emi * hter - hemi * ter - hemr * tei + mti * (hmm - htt));
            V[oMTI] = pmo * ((mm - tt) * hmtr + emr * hter - emi * htei + hemi * tei - hemr * ter - mtr * (hmm - htt));
        } else {   // the same terms in the order of nuosc_class.cpp:270-271
            V[oMTR] = pmo * ((tt - mm) * hmti + emr * htei + emi * hter - hemr * tei - hemi * ter + mti * (hmm - htt));
            V[oMTI] = pmo * ((mm - tt) * hmtr + emr * hter - emi * htei - hemr * ter + hemi * tei - mtr * (hmm - htt));
        }
    }
    // both tau-tau rows read the nu fields (nuosc_class.cpp:268-269)
    V[TT] = -(pmo2 * (hmti * c[MTR] + hmtr * c[MTI] + htei * c[TER] - hter * c[TEI]));
    V[BTT] = V[TT];
}

// Orders the LDS accesses of the block (LDS-only fences around the
// barrier, so outstanding global loads are not waited for).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Fields staged in LDS at a time: the x rows of the block plus 2 edge rows
// on each side, 64 directions: 36.9 KB per block in both precisions.
template <typename T>
constexpr int kGrp3 = sizeof(T) == 4 ? 18 : 9;

template <typename T>
struct XSlab {
    // rows i0-2 .. i0+R+1, then R-4 slots that waves 4.. write unread
    T x[kGrp3<T>][2 * kF3Rows][64];
};

// The stage input is never one of the stage's outputs (yin differs from u, w
// and yout in every stage): declared restrict, the stencil loads of later
// fields can be issued before the stores of earlier ones.
//
// x neighbours come from LDS: every wave writes its own row's centre values
// and one of the 4 rows beyond the block's edges (wave 0: i0-2, 1: i0-1,
// 2: i0+4, 3: i0+5), so each field costs 1 global load for the x stencil
// instead of 4.  All waves of the block reach every barrier; waves whose row
// is past nx load a stored row (at most nx + 1, a periodic ghost row) and
// compute nothing.
template <typename T, int STAGE, bool ONE>
__device__ __forceinline__ void stage3_body(const F3Args<T>& a, const T* __restrict__ yin, XSlab<T>& sx) {
    constexpr int kGrp = kGrp3<T>;
    const Geometry& g = a.g;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    // logical block -> (direction chunk, z tile, x quad, z site in the tile)
    const int jj = lb % kF3Tile;
    int r = lb / kF3Tile;
    const int qx = r % a.quads;
    r /= a.quads;
    const int jt = r % a.tiles;
    const int chg = r / a.tiles;
    const int j = jt * kF3Tile + jj;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // one x row (nx = 1): the x neighbours of every point are the point
    // itself (its periodic ghost rows are copies of it), so the waves of a
    // block take kF3Rows consecutive direction chunks of the row instead of
    // kF3Rows rows of one chunk; the x terms are still evaluated, from the
    // centre value, and the LDS steps below run unused
    constexpr bool one = ONE;
    const int ch = one ? chg * kF3Rows + w : chg;
    const int i = one ? 0 : qx * kF3Rows + w;
    const int lane = (int)(threadIdx.x & 63);
    if (j >= g.nz) return;   // whole blocks
    const bool row_ok = i < g.nx;
    const int il = min(i, g.nx + 1);                                  // row loaded
    // edge rows i0-2, i0-1, i0+R, i0+R+1 (LDS slots 0, 1, R+2, R+3), one for
    // each of waves 0-3; waves 4.. reload their own row into a slot nothing
    // reads (unconditional, so the compiler keeps the loads in flight)
    static_assert(kF3Rows >= 4, "four waves load the edge rows");
    const int er = w < 2 ? qx * kF3Rows - 2 + w : (w < 4 ? qx * kF3Rows + kF3Rows - 2 + w : il);
    const int el = min(er, g.nx + 1);
    const int eslot = w < 2 ? w : w + kF3Rows;
    // lanes past the last direction compute on the last one, store nothing
    // and add nothing to the moments (all lanes take part in the reduction)
    const bool live = ch * 64 + lane < g.nv;
    const int v = live ? ch * 64 + lane : g.nv - 1;

    const long long o = elem_off(g, il, j + g.zoff, 0) + v;
    const long long oe = elem_off(g, el, j + g.zoff, 0) + v;
    long long zo[4];   // z - 2, z - 1, z + 1, z + 2: stored ghost rows (open z) or the periodic wrap
    {
        const int ks[4] = {-2, -1, 1, 2};
#pragma unroll
        for (int q = 0; q < 4; ++q)
            zo[q] = (long long)(g.bc ? ks[q] : wrapi(j + ks[q], g.nz) - j) * g.srec;
    }

    // centre values and moment projections
    T c[kNVar3];
#pragma unroll
    for (int f = 0; f < kNVar3; ++f) c[f] = yin[o + (long long)f * g.nvp];
    const double vxd = g.vx[v], vyd = g.vy[v], vzd = g.vz[v];
    cdouble3* m = (cdouble3*)(a.mom + ((long long)min(i, g.nx - 1) * g.nz + j) * kNMom3);
    T P[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)   // INP, nuosc_class.cpp:187
        P[q] = (T)(m[4 * q] - m[4 * q + 1] * vxd - m[4 * q + 2] * vyd - m[4 * q + 3] * vzd);
    T h[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) h[q] = (T)a.h[q];
    T I[kNVar3], V[kNVar3];
    local_terms3<T>(c, P, (T)a.mu, (T)a.pmo, h, I, V);

    // advection and Kreiss-Oliger factors (nuosc_class.cpp:120-123, 180-183)
    const T fx = (T)(-vxd / g.twelve_dx), fz = (T)(-vzd / g.twelve_dz);
    const T kx = (T)g.ko_x, kz = (T)g.ko_z;
    const T cw = (T)a.cw, cy = (T)a.cy, ca = (T)a.ca, cb = (T)a.cb, third = (T)(1.0 / 3.0);
    // FP32: all loads and arithmetic first, the stores after, so no load
    // waits for a store (measured 5-10% faster); FP64: each field stored as
    // soon as it is done (holding all 18 results costs a wave per SIMD)
    constexpr bool kStoresLast = sizeof(T) == 4;
    T yo[kStoresLast ? kNVar3 : 1], wn[kStoresLast && STAGE == 2 ? kNVar3 : 1];
    double d[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // moment differences of the output (stages 1-7)
    const bool put = row_ok && live;
#pragma unroll
    for (int g0 = 0; g0 < kNVar3; g0 += kGrp) {
        if (g0 > 0) lds_barrier();   // every wave has read the previous group
#pragma unroll
        for (int q = 0; q < kGrp; ++q) {
            sx.x[q][w + 2][lane] = c[g0 + q];
            sx.x[q][eslot][lane] = yin[oe + (long long)(g0 + q) * g.nvp];
        }
        lds_barrier();
        if (row_ok) {
#pragma unroll
            for (int q = 0; q < kGrp; ++q) {
                const int f = g0 + q;
                const T* p = yin + o + (long long)f * g.nvp;
                const T zm2 = p[zo[0]], zm1 = p[zo[1]], zp1 = p[zo[2]], zp2 = p[zo[3]];
                const T xm2 = one ? c[f] : sx.x[q][w][lane], xm1 = one ? c[f] : sx.x[q][w + 1][lane];
                const T xp1 = one ? c[f] : sx.x[q][w + 3][lane], xp2 = one ? c[f] : sx.x[q][w + 4][lane];
                const T c0 = c[f];
                const T adv =
                    fz * ((zm2 - zp2) - (T)8.0 * (zm1 - zp1)) + fx * ((xm2 - xp2) - (T)8.0 * (xm1 - xp1));
                const T ko = kz * (zm2 + zp2 - (T)4 * (zm1 + zp1) + (T)6 * c0) +
                             kx * (xm2 + xp2 - (T)4 * (xm1 + xp1) + (T)6 * c0);
                const T k = I[f] + adv + ko + V[f];
                const long long of = o + (long long)f * g.nvp;
                T y, an = (T)0;
                if constexpr (STAGE == 0) {
                    y = k;
                } else if constexpr (STAGE == 1 || STAGE == 5) {   // Y2 = u + dt/2 k1; SSP-RK3 u1 = u + dt k
                    y = c0 + cy * k;
                } else if constexpr (STAGE == 2) {                 // A = Y2 + u + dt k2, Y3 = u + dt/2 k2
                    const T uc = a.u[of];
                    an = (c0 + uc) + cw * k;
                    y = uc + cy * k;
                } else if constexpr (STAGE == 3) {                 // Y4 = u + dt k3
                    y = a.u[of] + cy * k;
                } else if constexpr (STAGE == 4) {                 // u' = (A + Y4) / 3 + dt/6 k4
                    y = (a.w[of] + c0) * third + cw * k;
                } else {                                           // SSP-RK3 (nuosc_class.cpp:440, 444, 369)
                    y = ca * a.u[of] + cb * (c0 + cy * k);
                }
                if constexpr (STAGE != 0) {
                    d[kDiffOf[f][0]] += kDiffOf[f][1] > 0 ? (double)y : -(double)y;
                    if (kDiffOf[f][2] >= 0) d[kDiffOf[f][2]] += kDiffOf[f][3] > 0 ? (double)y : -(double)y;
                }
                if constexpr (kStoresLast) {
                    yo[f] = y;
                    if constexpr (STAGE == 2) wn[f] = an;
                } else if (live) {
                    if constexpr (STAGE == 2) a.w[of] = an;
                    a.yout[of] = y;
                }
            }
        }
    }
    if (!row_ok) return;   // after the last barrier
    if constexpr (kStoresLast) {
        if (put) {
#pragma unroll
            for (int f = 0; f < kNVar3; ++f) {
                const long long of = o + (long long)f * g.nvp;
                if constexpr (STAGE == 2) a.w[of] = wn[f];
                a.yout[of] = yo[f];
            }
        }
    }
    if constexpr (STAGE != 0) {
        // partial moments of the output over this wave's directions
        // (nuosc_class.cpp:65-71), reduced across the wave in a fixed order
        const double wv = live ? g.vw[v] : 0.0;
        const double q4[4] = {wv, vxd * wv, vyd * wv, vzd * wv};
        double t[32];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc)
#pragma unroll
            for (int q = 0; q < 4; ++q) t[4 * cc + q] = q4[q] * d[cc];
        const double tot = wave_sum32(t, lane);
        if ((lane & 1) == 0 && (!ONE || ch < a.nch))   // (one row: waves past the last chunk add nothing)
            a.part[(((long long)i * g.nz + j) * a.nch + ch) * kNMom3 + (lane >> 1)] = tot;
    }
}

// 4 waves per SIMD (128 VGPRs; FP64 stages 2-4, 6, 7 spill 2 registers):
// measured 7% faster per FP64 step than the 3 waves the compiler picks alone
template <typename T, int STAGE, bool ONE>
__global__ __launch_bounds__(64 * kF3Rows) __attribute__((amdgpu_waves_per_eu(4))) void k_stage3(const F3Args<T> a) {
    __shared__ XSlab<T> sx;
    stage3_body<T, STAGE, ONE>(a, a.yin, sx);
}

// ---------------------------------------------------------------------------
// k_stage3z: the 3-flavour stage as a z march (boxes with nx > 1).
//
// A block is 8 wavefronts over 4 adjacent x rows of one 64-direction chunk and
// one z segment.  Waves 0-3 (group A) own the fields ee, mm, emr, emi, bee,
// bmm, bemr, bemi, tt, btt of rows 0-3; waves 4-7 (group B) the fields mtr,
// mti, ter, tei, bmtr, bmti, bter, btei of the same rows.  The split follows
// the moment vectors: group A's outputs enter {emRm, emIp, eemm, mmtt}, group
// B's {mtRm, mtIp, teRm, teIp}, so each wave reduces 16 moment values and the
// two groups write disjoint parts of the (site, chunk) partial record that
// k_moment_totals3 adds.  Each wave keeps a 5-deep z window of its own fields
// in registers (every value of its row is loaded once per segment); the x
// neighbours, and the other group's centre values that the interaction and
// vacuum terms need, are read from a block-shared LDS slab: per z step every
// wave writes its row's centre values of its fields, and waves 0-3 (group A)
// and 4-7 (group B) load the 4 rows beyond the block's edges for their fields
// one step ahead.  The slab is double-buffered by the parity of the z step
// (FP64 147 KB: one block per CU, 2 waves per SIMD; FP32 74 KB: two blocks,
// 4 waves per SIMD), so one barrier per step orders both the writes before
// the reads and the reads before the rewrite two steps later (8% faster per
// FP64 step than a single slab with a second barrier).  u (or A) of the
// step is loaded after the interaction terms, right before the fields'
// loop: fewer registers live across the terms (FP32: 6-8 VGPRs spilled
// instead of 9-11; 2-10% faster per step in both precisions).
// ---------------------------------------------------------------------------
constexpr int kZW = 8;          // waves per block
constexpr int kZRows = 4;       // x rows per block
constexpr int kZSlots = kZRows + 4;   // rows -2 .. 5 of the block
constexpr int kNA = 10, kNB = 8;          // fields of group A and B
__device__ constexpr int kGroupField[2][kNA] = {{EE, MM, EMR, EMI, BEE, BMM, BEMR, BEMI, TT, BTT},
                                                {MTR, MTI, TER, TEI, BMTR, BMTI, BTER, BTEI, 0, 0}};
// record position of the group's moment value m (0..15): A holds d0, d1, d2,
// d7 (positions 0..11, 28..31), B holds d3..d6 (12..27)
__device__ __forceinline__ int rec_pos(int grp, int m) { return grp ? 12 + m : (m < 12 ? m : 16 + m); }

template <typename T>
struct ZSlab {
    T x[2][kNVar3][kZSlots][64];   // z steps alternate between the two buffers
};

typedef __attribute__((ext_vector_type(2))) unsigned int zu32x2;
constexpr int kZRsrcWord3 = 0x00020000;

// AUX: cache policy bits of the load (2: non-temporal)
template <typename T, int AUX = 0>
__device__ __forceinline__ T zld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, AUX));
    else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, AUX));
}
template <typename T>
__device__ __forceinline__ void zst(T v, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(zu32x2, v), r, vo, so, 2);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, 2);
}

template <typename T>
struct F3ZArgs {
    Geometry g;
    unsigned long long yin, u, w, yout;   // state copies (address of stored x row 0)
    const double* mom;                    // moment totals of yin: [site][32]
    double* part;                         // partial moments of yout: [site][nch][32]
    double cw, cy, ca, cb;
    double mu, pmo;
    double h[9];
    int quads, nch, zlen;
};

template <typename T, int STAGE, int GRP>
__device__ __forceinline__ void stage3z_body(const F3ZArgs<T>& a, ZSlab<T>& sl, int lb, int wr) {
    constexpr int grp = GRP;
    constexpr int NF = GRP ? kNB : kNA;   // fields of this wave
    const Geometry& g = a.g;
    const int lane = (int)(threadIdx.x & 63);
    // logical block -> (x quad fastest, direction chunk, z segment)
    const int qx = lb % a.quads;
    const int rest = lb / a.quads;
    const int ch = rest % a.nch;
    const int zs = rest / a.nch;
    const int nx = g.nx, nz = g.nz, nv = g.nv, nzs = g.nzs, zoff = g.zoff;
    const int ib = qx * kZRows;            // first x row of the block
    const int i = ib + wr;                 // this wave's x row
    const bool row_ok = i < nx;
    const int v = ch * 64 + lane;
    const bool live = v < nv;
    const int pv = min(v, nv - 1);         // lanes past nv work on the last direction and store nothing
    const unsigned esz = sizeof(T);
    const unsigned rowb = (unsigned)g.nvp * esz;
    const unsigned siteb = (unsigned)(g.srec * (long long)esz);
    const unsigned xrowb = (unsigned)(g.xrow * (long long)esz);
    // stored rows ib .. ib + kZRows + 2 kXg (ghost rows, spare row) under one
    // descriptor per copy; a lane's row relative to ib in its vector offset
    const int spare = nx + 2 * kXg - ib;
    const unsigned long long woff = (unsigned long long)ib * xrowb;
    const unsigned wbytes = (unsigned)(min(kZRows + 2 * kXg + 1, g.nxs - ib)) * xrowb;
    auto rs = [&](unsigned long long base) {
        unsigned long long b = base + woff;
        asm volatile("" : "+s"(b));
        return __builtin_amdgcn_make_buffer_rsrc((void*)b, 0, (int)wbytes, kZRsrcWord3);
    };
    const unsigned dof = (unsigned)pv * esz;
    const unsigned vo = (unsigned)min(wr + kXg, spare) * xrowb + dof;
    const unsigned vos = (unsigned)(row_ok && live ? wr + kXg : spare) * xrowb + dof;
    // edge rows: wave wr of each group loads block row -2, -1, 4, 5 (slots 0, 1, 6, 7)
    const int eb = wr < 2 ? wr - 2 : kZRows + wr - 2;
    const unsigned vh = (unsigned)min(eb + kXg, spare) * xrowb + dof;
    const int slot_h = eb + 2, slot_c = wr + 2;
    auto zrow = [&](int jj) { return (unsigned)wrapi(jj + zoff, nzs); };

    const double vxd = g.vx[pv], vyd = g.vy[pv], vzd = g.vz[pv];
    const T fx = (T)(-vxd / g.twelve_dx), fz = (T)(-vzd / g.twelve_dz);
    const T kx = (T)g.ko_x, kz = (T)g.ko_z;
    const T cw = (T)a.cw, cy = (T)a.cy, ca = (T)a.ca, cb = (T)a.cb, third = (T)(1.0 / 3.0);
    T h[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) h[q] = (T)a.h[q];
    const double wv = live ? g.vw[pv] : 0.0;
    const double q4[4] = {wv, vxd * wv, vyd * wv, vzd * wv};
    // the site's moment totals are wave-uniform: scalar loads
    typedef const __attribute__((address_space(4))) double cdz;
    cdz* mrow = (cdz*)(a.mom + (long long)min(i, nx - 1) * nz * kNMom3);

    const int j0 = zs * a.zlen;
    const int j1 = min(nz, j0 + a.zlen);
    T win[NF][5];   // own fields, z = j-2 .. j+2
    // Edge rows: step j loads the edge row of z+2, in the step in which the
    // neighbouring blocks load that row as their own, so it is mostly an L2
    // hit (HBM reads of the stage fell by 3-7 GB per FP64 launch against a
    // load one step later), and step j+1 writes it after its barrier into
    // the other buffer, the slab of step j+2.  FP64 issues the load right
    // after step j's barrier (latency hidden by the step's compute); FP32
    // at the end of step j, which keeps the row in registers only until the
    // next barrier (0-2 spilled VGPRs instead of 6-8: 14% faster per FP32
    // step; the same in FP64 was 4.5% slower).
    constexpr bool kEdgeAtEnd = sizeof(T) == 4;
    T hreg[NF];     // at the top of step j: the edge row of z+1
    {
        const __amdgpu_buffer_rsrc_t ry = rs(a.yin);
#pragma unroll
        for (int qz = 0; qz < 4; ++qz) {
            const unsigned so = zrow(j0 - 2 + qz) * siteb;
#pragma unroll
            for (int q = 0; q < NF; ++q)
                win[q][qz] = zld<T>(ry, vo, so + (unsigned)kGroupField[grp][q] * rowb);
        }
        // z = j0 straight into the slab of step j0, z = j0+1 into hreg
        const unsigned so = zrow(j0) * siteb, s1 = zrow(j0 + 1) * siteb;
#pragma unroll
        for (int q = 0; q < NF; ++q)
            sl.x[j0 & 1][kGroupField[grp][q]][slot_h][lane] = zld<T>(ry, vh, so + (unsigned)kGroupField[grp][q] * rowb);
#pragma unroll
        for (int q = 0; q < NF; ++q) hreg[q] = zld<T>(ry, vh, s1 + (unsigned)kGroupField[grp][q] * rowb);
    }
    for (int j = j0; j < j1; ++j) {
        const unsigned so = (unsigned)(j + zoff) * siteb;
        // ---- loads: own row z+2 (this step); the edge row follows the barrier
        const __amdgpu_buffer_rsrc_t ry = rs(a.yin);
        const unsigned s2 = zrow(j + 2) * siteb;
#pragma unroll
        for (int q = 0; q < NF; ++q)
            win[q][4] = zld<T>(ry, vo, s2 + (unsigned)kGroupField[grp][q] * rowb);
        __builtin_amdgcn_sched_barrier(0);
        // ---- LDS: own centre values of z (the edge rows of z are there)
        T (*X)[kZSlots][64] = sl.x[j & 1];
#pragma unroll
        for (int q = 0; q < NF; ++q) X[kGroupField[grp][q]][slot_c][lane] = win[q][2];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        {
            // every wave is done with step j-1, so the other buffer (the slab
            // of step j+1) takes the edge rows of z+1, published by the next
            // barrier; then (FP64) the edge rows of z+2 are loaded
            T (*Xn)[kZSlots][64] = sl.x[(j + 1) & 1];
#pragma unroll
            for (int q = 0; q < NF; ++q) Xn[kGroupField[grp][q]][slot_h][lane] = hreg[q];
            if constexpr (!kEdgeAtEnd) {
#pragma unroll
                for (int q = 0; q < NF; ++q)
                    hreg[q] = zld<T>(ry, vh, s2 + (unsigned)kGroupField[grp][q] * rowb);
            }
        }
        // ---- centre values of all 18 fields, moment projections, local terms
        T c[kNVar3];
#pragma unroll
        for (int f = 0; f < kNVar3; ++f) c[f] = X[f][slot_c][lane];
        T P[8];
        {
            const long long mo = (long long)j * kNMom3;
#pragma unroll
            for (int q = 0; q < 8; ++q)   // INP, nuosc_class.cpp:187
                P[q] = (T)(mrow[mo + 4 * q] - mrow[mo + 4 * q + 1] * vxd - mrow[mo + 4 * q + 2] * vyd -
                           mrow[mo + 4 * q + 3] * vzd);
        }
        T I[kNVar3], V[kNVar3];
        local_terms3<T>(c, P, (T)a.mu, (T)a.pmo, h, I, V);
        // (u and A are read once per stage: loaded with the non-temporal hint,
        // 1.5-1.8% faster per step in both precisions, medians of 4 runs)
        T uc[NF];
        if constexpr (STAGE == 2 || STAGE == 3 || STAGE == 6 || STAGE == 7 || STAGE == 4) {
            const __amdgpu_buffer_rsrc_t ru = rs(STAGE == 4 ? a.w : a.u);
#pragma unroll
            for (int q = 0; q < NF; ++q)
                uc[q] = zld<T, 2>(ru, vo, so + (unsigned)kGroupField[grp][q] * rowb);
        }
        const __amdgpu_buffer_rsrc_t ro = rs(a.yout);
        const __amdgpu_buffer_rsrc_t rw = rs(a.w);
        double t[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) t[s] = 0.0;
        double d[4] = {0.0, 0.0, 0.0, 0.0};   // group A: d0, d1, d2, d7; group B: d3..d6
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const int f = kGroupField[grp][q];
            const T c0 = win[q][2];
            const T xm2 = X[f][slot_c - 2][lane], xm1 = X[f][slot_c - 1][lane];
            const T xp1 = X[f][slot_c + 1][lane], xp2 = X[f][slot_c + 2][lane];
            const T zm2 = win[q][0], zm1 = win[q][1], zp1 = win[q][3], zp2 = win[q][4];
            // advection and Kreiss-Oliger (nuosc_class.cpp:120-123, 180-183)
            const T adv = fz * ((zm2 - zp2) - (T)8.0 * (zm1 - zp1)) + fx * ((xm2 - xp2) - (T)8.0 * (xm1 - xp1));
            const T ko = kz * (zm2 + zp2 - (T)4 * (zm1 + zp1) + (T)6 * c0) +
                         kx * (xm2 + xp2 - (T)4 * (xm1 + xp1) + (T)6 * c0);
            const T k = I[f] + adv + ko + V[f];
            T y, an = (T)0;
            if constexpr (STAGE == 0) {
                y = k;
            } else if constexpr (STAGE == 1 || STAGE == 5) {   // Y2 = u + dt/2 k1; SSP-RK3 u1 = u + dt k
                y = c0 + cy * k;
            } else if constexpr (STAGE == 2) {                 // A = Y2 + u + dt k2, Y3 = u + dt/2 k2
                an = (c0 + uc[q]) + cw * k;
                y = uc[q] + cy * k;
            } else if constexpr (STAGE == 3) {                 // Y4 = u + dt k3
                y = uc[q] + cy * k;
            } else if constexpr (STAGE == 4) {                 // u' = (A + Y4) / 3 + dt/6 k4
                y = (uc[q] + c0) * third + cw * k;
            } else {                                           // SSP-RK3 (nuosc_class.cpp:440, 444, 369)
                y = ca * uc[q] + cb * (c0 + cy * k);
            }
            {
                const unsigned fo = so + (unsigned)f * rowb;
                if constexpr (STAGE == 2) zst<T>(an, rw, vos, fo);
                zst<T>(y, ro, vos, fo);
            }
            if constexpr (STAGE != 0) {
                // this field's share of the moment differences (kDiffOf), in
                // the group's d slots: A {0,1,2,7} -> 0..3, B {3..6} -> 0..3
                const int d0 = kDiffOf[f][0], d1 = kDiffOf[f][2];
                const int s0 = grp ? d0 - 3 : (d0 == 7 ? 3 : d0);
                d[s0] += kDiffOf[f][1] > 0 ? (double)y : -(double)y;
                if (d1 >= 0) {
                    const int s1 = grp ? d1 - 3 : (d1 == 7 ? 3 : d1);
                    d[s1] += kDiffOf[f][3] > 0 ? (double)y : -(double)y;
                }
            }
        }
        if constexpr (STAGE != 0) {
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
#pragma unroll
                for (int s = 0; s < 4; ++s) t[4 * cc + s] = q4[s] * d[cc];
            // wave total of the 16 values: lane L ends with value L >> 2
            const double tot = wave_sum16(t);
            if (row_ok && (lane & 3) == 0)
                a.part[(((long long)i * nz + j) * a.nch + ch) * kNMom3 + rec_pos(grp, lane >> 2)] = tot;
        }
        if constexpr (kEdgeAtEnd) {   // (FP32) the edge rows of z+2, written after the next barrier
#pragma unroll
            for (int q = 0; q < NF; ++q) hreg[q] = zld<T>(ry, vh, s2 + (unsigned)kGroupField[grp][q] * rowb);
        }
        // no second barrier: step j+1 writes the other buffer, and step j+2's
        // writes to this one follow step j+1's barrier, which every wave
        // reaches after its reads of step j
#pragma unroll
        for (int q = 0; q < NF; ++q)
#pragma unroll
            for (int qz = 0; qz < 4; ++qz) win[q][qz] = win[q][qz + 1];
    }
}

// waves 0-3 run group A's fields, waves 4-7 group B's: both execute the same
// number of barriers per z step, which is what the block barrier counts
template <typename T, int STAGE>
__global__ __launch_bounds__(64 * kZW) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 2 : 4)))
void k_stage3z(const F3ZArgs<T> a) {
    __shared__ ZSlab<T> sl;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wib < 4) stage3z_body<T, STAGE, 0>(a, sl, lb, wib);
    else stage3z_body<T, STAGE, 1>(a, sl, lb, wib - 4);
}

// per-site totals of the partial records, summed in chunk order
__global__ __launch_bounds__(256) void k_moment_totals3(const double* __restrict__ part, int nch, long long n,
                                                        double* __restrict__ mom) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // site * 32 + value
    if (k >= n) return;
    const long long site = k / kNMom3, q = k % kNMom3;
    const double* p = part + site * nch * kNMom3 + q;
    double s = p[0];
    for (int c = 1; c < nch; ++c) s += p[(long long)c * kNMom3];
    mom[k] = s;
}

// k_stage3 runs one-row grids only (ONE = true): boxes take the z march, and
// boxes it cannot address are refused at setup (flavor3_addressable).  The
// one-row form keeps the multi-row kernel's LDS staging and barriers (their
// values unused): a copy without them, and one with scheduling barriers in
// their place, measured 8-11% slower per 1x4096x642 column step.
template <typename T, int STAGE>
hipError_t launch3(const F3Args<T>& a, unsigned nb, hipStream_t s) {
    if (!a.one_row) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_stage3<T, STAGE, true>), dim3(nb), dim3(64 * kF3Rows), 0, s, a);
    return hipGetLastError();
}

template <typename T, int STAGE>
hipError_t launch3z(const F3ZArgs<T>& a, unsigned nb, hipStream_t s) {
    hipLaunchKernelGGL((k_stage3z<T, STAGE>), dim3(nb), dim3(64 * kZW), 0, s, a);
    return hipGetLastError();
}

}  // namespace

// The z march (boxes with nx > 1) addresses the x rows of one block (4 rows,
// 3 ghost rows each side, the spare row and one of margin) through 32-bit
// buffer offsets: they must span less than 4 GiB.
bool flavor3_addressable(const Geometry& g, int esize) {
    const long long span = (long long)(kZRows + 2 * kXg + 2) * g.xrow * esize;
    return g.nx == 1 || span < (1ll << 32) - 1;
}

template <typename T>
hipError_t launch_moments3(const Geometry& g, const T* y, double* mom, hipStream_t s) {
    const long long sites = (long long)g.nx * g.nz;
    if (g.nf != kNVar3 || sites <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_moments3<T>), dim3((unsigned)((sites + 3) / 4)), dim3(256), 0, s, g, y, mom);
    return hipGetLastError();
}

hipError_t launch_moment_totals3(const Geometry& g, const double* part, double* mom, hipStream_t s) {
    const long long n = (long long)g.nx * g.nz * kNMom3;
    if (g.nf != kNVar3 || n <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_moment_totals3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part,
                       (g.nv + 63) / 64, n, mom);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_stage3(int stage, const Geometry& g, const Physics& ph, const T* yin, const T* u, T* w, T* yout,
                         const double* mom, double* part, double cw, double cy, hipStream_t s, double ca,
                         double cb) {
    if (g.nf != kNVar3 || (stage != 0 && !part)) return hipErrorInvalidValue;
    if (g.nx > 1) {
        if (!flavor3_addressable(g, (int)sizeof(T))) return hipErrorInvalidValue;
        F3ZArgs<T> z{};
        z.g = g;
        z.yin = (unsigned long long)yin;
        z.u = (unsigned long long)(u ? u : yin);
        z.w = (unsigned long long)(w ? w : yout);
        z.yout = (unsigned long long)yout;
        z.mom = mom;
        z.part = part;
        z.cw = cw;
        z.cy = cy;
        z.ca = ca;
        z.cb = cb;
        z.mu = ph.mu;
        z.pmo = ph.pmo;
        for (int q = 0; q < 9; ++q) z.h[q] = ph.h3[q];
        z.quads = (g.nx + kZRows - 1) / kZRows;
        z.nch = (g.nv + 63) / 64;
        // z segments of 256 sites, shorter while that leaves fewer than 4
        // blocks per CU's worth (256 CUs) of work (segments of 128-256 sites
        // measured 8% faster per FP64 step than 64 at 512 x 512 x 642)
        int zl = 256;
        while (zl > 8 && (long long)z.quads * z.nch * ((g.nz + zl - 1) / zl) < 4 * 256) zl /= 2;
        z.zlen = std::min(zl, g.nz);
        const long long nb = (long long)z.quads * z.nch * ((g.nz + z.zlen - 1) / z.zlen);
        if (nb <= 0 || nb > 0x7fffffffLL) return hipErrorInvalidValue;
        switch (stage) {
            case 0: return launch3z<T, 0>(z, (unsigned)nb, s);
            case 1: return launch3z<T, 1>(z, (unsigned)nb, s);
            case 2: return launch3z<T, 2>(z, (unsigned)nb, s);
            case 3: return launch3z<T, 3>(z, (unsigned)nb, s);
            case 4: return launch3z<T, 4>(z, (unsigned)nb, s);
            case 5: return launch3z<T, 5>(z, (unsigned)nb, s);
            case 6: return launch3z<T, 6>(z, (unsigned)nb, s);
            case 7: return launch3z<T, 7>(z, (unsigned)nb, s);
            default: return hipErrorInvalidValue;
        }
    }
    F3Args<T> a{};
    a.g = g;
    a.yin = yin;
    a.u = u;
    a.w = w;
    a.yout = yout;
    a.mom = mom;
    a.part = part;
    a.nch = (g.nv + 63) / 64;
    a.cw = cw;
    a.cy = cy;
    a.ca = ca;
    a.cb = cb;
    a.mu = ph.mu;
    a.pmo = ph.pmo;
    for (int q = 0; q < 9; ++q) a.h[q] = ph.h3[q];
    a.tiles = (g.nz + kF3Tile - 1) / kF3Tile;
    a.quads = (g.nx + kF3Rows - 1) / kF3Rows;
    a.one_row = g.nx == 1;
    a.cgroups = a.one_row ? (a.nch + kF3Rows - 1) / kF3Rows : a.nch;
    const long long nb = (long long)a.cgroups * a.tiles * a.quads * kF3Tile;
    if (nb <= 0 || nb > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (stage) {
        case 0: return launch3<T, 0>(a, (unsigned)nb, s);
        case 1: return launch3<T, 1>(a, (unsigned)nb, s);
        case 2: return launch3<T, 2>(a, (unsigned)nb, s);
        case 3: return launch3<T, 3>(a, (unsigned)nb, s);
        case 4: return launch3<T, 4>(a, (unsigned)nb, s);
        case 5: return launch3<T, 5>(a, (unsigned)nb, s);
        case 6: return launch3<T, 6>(a, (unsigned)nb, s);
        case 7: return launch3<T, 7>(a, (unsigned)nb, s);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_moments3<double>(const Geometry&, const double*, double*, hipStream_t);
template hipError_t launch_moments3<float>(const Geometry&, const float*, double*, hipStream_t);
template hipError_t launch_stage3<double>(int, const Geometry&, const Physics&, const double*, const double*,
                                          double*, double*, const double*, double*, double, double, hipStream_t,
                                          double, double);
template hipError_t launch_stage3<float>(int, const Geometry&, const Physics&, const float*, const float*, float*,
                                         float*, const double*, double*, double, double, hipStream_t, double,
                                         double);

}  // namespace nuosc
