// nuosc_kernels.hip -- CDNA4 (gfx950) kernels of the nuosc RK4 time stepper.
//
// The hot path is k_stage: one launch per RK4 stage.  Each wavefront owns one
// x-row i, one 64-direction chunk of v (lane = direction) and one z-segment,
// and marches along z keeping a 5-point z-window of all 8 fields in registers
// (every value of the row is loaded once).  The x-neighbours i-2..i+2 are
// plain coalesced loads that hit in L2: a block->tile remap keeps the waves of
// adjacent x-rows of the same (z-segment, v-chunk) on one XCD at the same
// time.  The velocity moments needed by the interaction term (a reduction
// over ALL directions of a site) are produced by the previous stage's
// epilogue as per-(site, chunk) partial sums, so no separate pass over the
// state is needed; the consumer adds the partials of a site in chunk order.
//
// RK4 is streamed with four state copies u, w, yA, yB (w accumulates the
// final combination), i.e. 16 passes over the 8 fields per step:
//   S1: read u            -> w = u + dt/6 k1,  yA = u + dt/2 k1
//   S2: read yA, u, w     -> w += dt/3 k2,     yB = u + dt/2 k2
//   S3: read yB, u, w     -> w += dt/3 k3,     yA = u + dt k3
//   S4: read yA, w        -> u  = w + dt/6 k4
// which equals the reference's combination u + dt/6 (k4 + (k1 + 2k2) + 2k3)
// (nuosc_class.cpp:374-427) up to rounding.
//
// The right-hand side follows nuosc_class.cpp:11-307 (2 flavours, 3rd-order
// Kreiss-Oliger, 4th-order central advection, vacuum term), with the
// reference's expression order.
#include <hip/hip_runtime.h>

#include "nuosc_dev.h"

#include <cstdint>

namespace nuosc {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrapi(int k, int n) {
    k %= n;
    return k < 0 ? k + n : k;
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Return the logical block index so that each XCD works on one contiguous
// range of logical blocks (a bijection for any grid size).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int per = nb / 8, rem = nb % 8;
    const int x = b % 8, loc = b / 8;
    return (x < rem) ? x * (per + 1) + loc : rem * (per + 1) + (x - rem) * per + loc;
}

// Knuth/Dekker two-sum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Wave-wide sums of 16 values with 17 cross-lane exchanges: after the call
// lane 4*m + r (r = 0..3) holds the total of a[m] over the 64 lanes.
template <bool COMP>
__device__ __forceinline__ double wave_sum16(const double (&a)[16], int lane) {
    if constexpr (!COMP) {
        double b[8], c[4], d[2];
        const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double send = h5 ? a[k] : a[k + 8];
            const double keep = h5 ? a[k + 8] : a[k];
            b[k] = keep + __shfl_xor(send, 32);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double send = h4 ? b[k] : b[k + 4];
            const double keep = h4 ? b[k + 4] : b[k];
            c[k] = keep + __shfl_xor(send, 16);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double send = h3 ? c[k] : c[k + 2];
            const double keep = h3 ? c[k + 2] : c[k];
            d[k] = keep + __shfl_xor(send, 8);
        }
        double e = (h2 ? d[1] : d[0]) + __shfl_xor(h2 ? d[0] : d[1], 4);
        e += __shfl_xor(e, 2);
        e += __shfl_xor(e, 1);
        return e;
    } else {
        // same exchange pattern carrying (sum, error) pairs
        double bs[8], be[8], cs[4], ce[4], ds[2], de[2];
        const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double send = h5 ? a[k] : a[k + 8];
            const double keep = h5 ? a[k + 8] : a[k];
            two_sum(keep, __shfl_xor(send, 32), bs[k], be[k]);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double ss = h4 ? bs[k] : bs[k + 4], se = h4 ? be[k] : be[k + 4];
            const double ks = h4 ? bs[k + 4] : bs[k], ke = h4 ? be[k + 4] : be[k];
            double err;
            two_sum(ks, __shfl_xor(ss, 16), cs[k], err);
            ce[k] = ke + __shfl_xor(se, 16) + err;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double ss = h3 ? cs[k] : cs[k + 2], se = h3 ? ce[k] : ce[k + 2];
            const double ks = h3 ? cs[k + 2] : cs[k], ke = h3 ? ce[k + 2] : ce[k];
            double err;
            two_sum(ks, __shfl_xor(ss, 8), ds[k], err);
            de[k] = ke + __shfl_xor(se, 8) + err;
        }
        double es, ee, err;
        {
            const double ss = h2 ? ds[0] : ds[1], se = h2 ? de[0] : de[1];
            const double ks = h2 ? ds[1] : ds[0], ke = h2 ? de[1] : de[0];
            two_sum(ks, __shfl_xor(ss, 4), es, err);
            ee = ke + __shfl_xor(se, 4) + err;
        }
#pragma unroll
        for (int m = 2; m >= 1; m >>= 1) {
            double s2;
            two_sum(es, __shfl_xor(es, m), s2, err);
            ee = ee + __shfl_xor(ee, m) + err;
            es = s2;
        }
        return es + ee;
    }
}

// Per-lane moment terms of one point: q = (w, vx w, vy w, vz w) times the
// three differences (nuosc_class.cpp:65-71).
template <typename T>
__device__ __forceinline__ void moment_terms(const T (&y)[kNVar], const double (&q)[4],
                                             double (&t)[16]) {
    const double d1 = (double)y[6] - (double)y[2];                                   // bemr - emr
    const double d2 = (double)y[7] + (double)y[3];                                   // bemi + emi
    const double d3 = (((double)y[4] - (double)y[5]) - (double)y[0]) + (double)y[1]; // bee-bmm-ee+mm
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        t[s] = q[s] * d1;
        t[4 + s] = q[s] * d2;
        t[8 + s] = q[s] * d3;
    }
#pragma unroll
    for (int s = 12; s < 16; ++s) t[s] = 0.0;
}

// Sum of the nch partial moment records of one site, in chunk order.
template <bool COMP>
__device__ __forceinline__ void site_moments(const double* __restrict__ rec, int nch,
                                             double (&M)[kNMom]) {
    if constexpr (!COMP) {
#pragma unroll
        for (int s = 0; s < kNMom; ++s) M[s] = rec[s];
        for (int c = 1; c < nch; ++c) {
#pragma unroll
            for (int s = 0; s < kNMom; ++s) M[s] += rec[c * kMomStride + s];
        }
    } else {
        double E[kNMom];
#pragma unroll
        for (int s = 0; s < kNMom; ++s) { M[s] = rec[s]; E[s] = 0.0; }
        for (int c = 1; c < nch; ++c) {
#pragma unroll
            for (int s = 0; s < kNMom; ++s) {
                double sum, err;
                two_sum(M[s], rec[c * kMomStride + s], sum, err);
                M[s] = sum;
                E[s] += err;
            }
        }
#pragma unroll
        for (int s = 0; s < kNMom; ++s) M[s] += E[s];
    }
}

// ---------------------------------------------------------------------------
// RK4 stage kernel
// ---------------------------------------------------------------------------
template <typename T>
struct StageArgs {
    Geometry g;
    const T* __restrict__ yin;
    const T* __restrict__ u;
    T* __restrict__ w;
    T* __restrict__ yout;
    const double* __restrict__ min;
    double* __restrict__ mout;
    double cw, cy;
    double mu2, mu, p2st, p2ct, pmo, ct2, st;
    int zlen, ntask;
};

template <typename T, int STAGE, bool COMP>
__global__ __launch_bounds__(256) void k_stage(const StageArgs<T> a) {
    const int lane = threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int task = xcd_remap(blockIdx.x, gridDim.x) * 4 + wib;
    if (task >= a.ntask) return;

    const int nx = a.g.nx, nz = a.g.nz, nvp = a.g.nvp, nch = a.g.nch, nv = a.g.nv;
    const long long fs = a.g.fs;
    const int i = task % nx;
    const int rest = task / nx;
    const int ch = rest % nch;
    const int zs = rest / nch;
    const int v = ch * kWave + lane;
    const bool act = v < nv;
    const int vl = act ? v : nv - 1;

    // per-direction constants (nuosc_class.cpp:119-121, 180-181)
    const double vxd = a.g.vx[vl], vyd = a.g.vy[vl], vzd = a.g.vz[vl];
    const double vwd = act ? a.g.vw[vl] : 0.0;
    const T fx = (T)(-vxd / a.g.twelve_dx);
    const T fz = (T)(-vzd / a.g.twelve_dz);
    const T kx = (T)a.g.ko_x, kz = (T)a.g.ko_z;
    const double q[4] = {vwd, vxd * vwd, vyd * vwd, vzd * vwd};
    const T mu2 = (T)a.mu2, mu = (T)a.mu, p2st = (T)a.p2st, p2ct = (T)a.p2ct;
    const T pmo = (T)a.pmo, ct2 = (T)a.ct2, st = (T)a.st;
    const T cw = (T)a.cw, cy = (T)a.cy;

    // row bases (site index of (i', 0))
    const long long r0 = (long long)i * nz;
    const long long rm1 = (long long)wrapi(i - 1, nx) * nz;
    const long long rm2 = (long long)wrapi(i - 2, nx) * nz;
    const long long rp1 = (long long)wrapi(i + 1, nx) * nz;
    const long long rp2 = (long long)wrapi(i + 2, nx) * nz;

    const T* __restrict__ yin = a.yin;
    auto ld = [&](const T* __restrict__ base, int f, long long site) -> T {
        return base[f * fs + site * nvp + vl];
    };

    const int j0 = zs * a.zlen;
    const int j1 = min(nz, j0 + a.zlen);

    // z-window: win[f][0..4] = field f at j-2 .. j+2 of row i
    T win[kNVar][5];
#pragma unroll
    for (int qz = 0; qz < 4; ++qz) {
        const long long site = r0 + wrapi(j0 - 2 + qz, nz);
#pragma unroll
        for (int f = 0; f < kNVar; ++f) win[f][qz] = ld(yin, f, site);
    }

    for (int j = j0; j < j1; ++j) {
        const long long site = r0 + j;
        {
            const long long s2 = r0 + wrapi(j + 2, nz);
#pragma unroll
            for (int f = 0; f < kNVar; ++f) win[f][4] = ld(yin, f, s2);
        }
        // velocity moments of the stage input at this site
        double M[kNMom];
        site_moments<COMP>(a.min + site * nch * kMomStride, nch, M);
        // projections onto direction v (INP, nuosc_class.cpp:187)
        const T A = (T)(M[0] - M[1] * vxd - M[2] * vyd - M[3] * vzd);
        const T B = (T)(M[4] - M[5] * vxd - M[6] * vyd - M[7] * vzd);
        const T C = (T)(M[8] - M[9] * vxd - M[10] * vyd - M[11] * vzd);

        const T ee = win[0][2], mm = win[1][2], emr = win[2][2], emi = win[3][2];
        const T bee = win[4][2], bmm = win[5][2], bemr = win[6][2], bemi = win[7][2];
        // interaction (nuosc_class.cpp:193-200)
        const T Iee = mu2 * (emi * A + emr * B);
        const T Iemr = -mu * ((ee - mm) * B + C * emi);
        const T Iemi = mu * ((mm - ee) * A + C * emr);
        const T Ibee = mu2 * (bemi * A + bemr * B);
        const T Ibemr = -mu * ((bee - bmm) * B + C * bemi);
        const T Ibemi = mu * ((bmm - bee) * A + C * bemr);
        // local terms (interaction + vacuum, nuosc_class.cpp:280-287)
        T loc[kNVar];
        loc[0] = Iee;
        loc[1] = -Iee;
        loc[2] = Iemr;
        loc[3] = Iemi;
        loc[4] = Ibee;
        loc[5] = -Ibee;
        loc[6] = Ibemr;
        loc[7] = Ibemi;
        T vac[kNVar];
        vac[0] = -(p2st * emi);
        vac[1] = p2st * emi;
        vac[2] = -(p2ct * emi);
        vac[3] = pmo * (ct2 * emr + st * (ee - mm));
        vac[4] = -(p2st * bemi);
        vac[5] = p2st * bemi;
        vac[6] = -(p2ct * bemi);
        vac[7] = pmo * (ct2 * bemr + st * (bee - bmm));

        T k[kNVar];
#pragma unroll
        for (int f = 0; f < kNVar; ++f) {
            const T xm2 = ld(yin, f, rm2 + j), xm1 = ld(yin, f, rm1 + j);
            const T xp1 = ld(yin, f, rp1 + j), xp2 = ld(yin, f, rp2 + j);
            const T zm2 = win[f][0], zm1 = win[f][1], c0 = win[f][2], zp1 = win[f][3], zp2 = win[f][4];
            // advection (nuosc_class.cpp:182-183)
            const T adv = fz * ((zm2 - zp2) - (T)8.0 * (zm1 - zp1)) + fx * ((xm2 - xp2) - (T)8.0 * (xm1 - xp1));
            // Kreiss-Oliger (nuosc_class.cpp:122-123)
            const T ko = kz * (zm2 + zp2 - (T)4 * (zm1 + zp1) + (T)6 * c0) +
                         kx * (xm2 + xp2 - (T)4 * (xm1 + xp1) + (T)6 * c0);
            k[f] = loc[f] + adv + ko + vac[f];
        }

        T yo[kNVar];
        if constexpr (STAGE == 0) {
#pragma unroll
            for (int f = 0; f < kNVar; ++f) yo[f] = k[f];
        } else if constexpr (STAGE == 1) {
#pragma unroll
            for (int f = 0; f < kNVar; ++f) {
                const T uc = win[f][2];
                const T wn = uc + cw * k[f];
                yo[f] = uc + cy * k[f];
                if (act) a.w[f * fs + site * nvp + v] = wn;
            }
        } else if constexpr (STAGE == 2 || STAGE == 3) {
#pragma unroll
            for (int f = 0; f < kNVar; ++f) {
                const T uc = ld(a.u, f, site);
                const T wc = ld(a.w, f, site);
                const T wn = wc + cw * k[f];
                yo[f] = uc + cy * k[f];
                if (act) a.w[f * fs + site * nvp + v] = wn;
            }
        } else {
#pragma unroll
            for (int f = 0; f < kNVar; ++f) {
                const T wc = ld(a.w, f, site);
                yo[f] = wc + cw * k[f];
            }
        }
        if (act) {
#pragma unroll
            for (int f = 0; f < kNVar; ++f) a.yout[f * fs + site * nvp + v] = yo[f];
        }
        if constexpr (STAGE != 0) {
            double t[16];
            moment_terms<T>(yo, q, t);
            const double tot = wave_sum16<COMP>(t, lane);
            const int m = lane >> 2;
            if ((lane & 3) == 0 && m < kNMom) a.mout[(site * nch + ch) * kMomStride + m] = tot;
        }
#pragma unroll
        for (int f = 0; f < kNVar; ++f) {
            win[f][0] = win[f][1];
            win[f][1] = win[f][2];
            win[f][2] = win[f][3];
            win[f][3] = win[f][4];
        }
    }
}

StagePlan plan_stage(const Geometry& g, int num_cus) {
    // enough waves to give every SIMD several (>= 16 per CU), segments long
    // enough that the 4-row window warm-up stays a small overhead
    StagePlan p{};
    const long long rows = (long long)g.nx * g.nch;
    const long long want = (long long)num_cus * 48;
    int zlen = g.nz;
    while (zlen > 8 && rows * ((g.nz + zlen - 1) / zlen) < want) zlen = (zlen + 1) / 2;
    if (zlen > 64 && rows * ((g.nz + 63) / 64) >= want) zlen = 64;
    p.zlen = zlen;
    p.nzs = (g.nz + zlen - 1) / zlen;
    p.ntask = (int)(rows * p.nzs);
    p.nblocks = (p.ntask + 3) / 4;
    return p;
}

template <typename T>
hipError_t launch_stage(int stage, const Geometry& g, const Physics& ph, const StagePlan& pl,
                        const T* yin, const T* u, T* w, T* yout, const double* min, double* mout,
                        double cw, double cy, hipStream_t s) {
    StageArgs<T> a;
    a.g = g;
    a.yin = yin;
    a.u = u;
    a.w = w;
    a.yout = yout;
    a.min = min;
    a.mout = mout;
    a.cw = cw;
    a.cy = cy;
    a.mu2 = 2 * ph.mu;
    a.mu = ph.mu;
    a.p2st = ph.pmo * 2 * ph.st;
    a.p2ct = ph.pmo * 2 * ph.ct;
    a.pmo = ph.pmo;
    a.ct2 = 2 * ph.ct;
    a.st = ph.st;
    a.zlen = pl.zlen;
    a.ntask = pl.ntask;
    constexpr bool COMP = sizeof(T) == 4;
    dim3 grid(pl.nblocks), block(256);
    switch (stage) {
        case 0: hipLaunchKernelGGL((k_stage<T, 0, COMP>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_stage<T, 1, COMP>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_stage<T, 2, COMP>), grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_stage<T, 3, COMP>), grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_stage<T, 4, COMP>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// stand-alone moment pass (after init / upload): one wave per (site, chunk)
// ---------------------------------------------------------------------------
template <typename T, bool COMP>
__global__ __launch_bounds__(256) void k_moments(const Geometry g, const T* __restrict__ y,
                                                 double* __restrict__ mout, long long ntask) {
    const int lane = threadIdx.x & 63;
    const long long task = (long long)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (task >= ntask) return;
    const long long site = task / g.nch;
    const int ch = (int)(task % g.nch);
    const int v = ch * kWave + lane;
    const bool act = v < g.nv;
    const int vl = act ? v : g.nv - 1;
    const double vwd = act ? g.vw[vl] : 0.0;
    const double q[4] = {vwd, g.vx[vl] * vwd, g.vy[vl] * vwd, g.vz[vl] * vwd};
    T yv[kNVar];
#pragma unroll
    for (int f = 0; f < kNVar; ++f) yv[f] = y[f * g.fs + site * g.nvp + vl];
    double t[16];
    moment_terms<T>(yv, q, t);
    const double tot = wave_sum16<COMP>(t, lane);
    const int m = lane >> 2;
    if ((lane & 3) == 0 && m < kNMom) mout[(site * g.nch + ch) * kMomStride + m] = tot;
}

template <typename T>
hipError_t launch_moments(const Geometry& g, const T* y, double* mout, hipStream_t s) {
    const long long ntask = (long long)g.nx * g.nz * g.nch;
    const long long nb = (ntask + 3) / 4;
    hipLaunchKernelGGL((k_moments<T, sizeof(T) == 4>), dim3((unsigned)nb), dim3(256), 0, s, g, y, mout,
                       ntask);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// initial data (nuosc_init.cpp:74-118 and the seeded complex perturbation)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

template <typename T>
__global__ __launch_bounds__(256) void k_init(const Geometry g, int ipt, const double* __restrict__ G0,
                                              const double* __restrict__ G0b,
                                              const double* __restrict__ tmpr_z,
                                              const double* __restrict__ p3o_z, double eps0,
                                              uint64_t seed, T* __restrict__ y) {
#pragma clang fp contract(off)
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.nz * g.nvp;
    if (n >= total) return;
    const long long site = n / g.nvp;
    const int v = (int)(n % g.nvp);
    if (v >= g.nv) {  // row padding
#pragma unroll
        for (int f = 0; f < kNVar; ++f) y[f * g.fs + n] = (T)0;
        return;
    }
    const int j = (int)(site % g.nz);
    double tmpr, tmpi = 0.0, p3o;
    if (ipt == 5) {
        const uint64_t key = splitmix64(seed ^ 0x6A09E667F3BCC909ull);
        const uint64_t lin = (uint64_t)(site * g.nv + v);
        const uint64_t h1 = splitmix64(key + 2ull * lin);
        const uint64_t h2 = splitmix64(key + 2ull * lin + 1ull);
        const double ur = 2.0 * ((double)(h1 >> 11) * 0x1.0p-53) - 1.0;
        const double ui = 2.0 * ((double)(h2 >> 11) * 0x1.0p-53) - 1.0;
        tmpr = eps0 * ur;
        tmpi = eps0 * ui;
        p3o = __builtin_sqrt(1.0 - tmpr * tmpr - tmpi * tmpi);
    } else {
        tmpr = tmpr_z[j];
        p3o = p3o_z[j];
    }
    const double hg = 0.5 * G0[v], hgb = 0.5 * G0b[v];
    y[0 * g.fs + n] = (T)(hg * (1.0 + p3o));
    y[1 * g.fs + n] = (T)(hg * (1.0 - p3o));
    y[2 * g.fs + n] = (T)(hg * tmpr);
    y[3 * g.fs + n] = (T)(ipt == 5 ? hg * tmpi : 0.0);
    y[4 * g.fs + n] = (T)(hgb * (1.0 + p3o));
    y[5 * g.fs + n] = (T)(hgb * (1.0 - p3o));
    y[6 * g.fs + n] = (T)(hgb * tmpr);
    y[7 * g.fs + n] = (T)(ipt == 5 ? -hgb * tmpi : 0.0);
}

template <typename T>
hipError_t launch_init(const Geometry& g, int ipt, const double* G0, const double* G0b,
                       const double* tmpr_z, const double* p3o_z, double eps0, uint64_t seed, T* y,
                       hipStream_t s) {
    const long long total = (long long)g.nx * g.nz * g.nvp;
    hipLaunchKernelGGL((k_init<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, ipt, G0,
                       G0b, tmpr_z, p3o_z, eps0, seed, y);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// diagnostics (nuosc_ana.cpp:3-137): compensated FP64 sums, fixed order
// ---------------------------------------------------------------------------
constexpr int kAnaThreads = 256;

int ana_blocks(const Geometry& g) {
    const long long pts = (long long)g.nx * g.nz * g.nv;
    long long nb = (pts + kAnaThreads * 16 - 1) / (kAnaThreads * 16);
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    return (int)nb;
}

template <typename T>
__global__ __launch_bounds__(kAnaThreads) void k_analysis(const Geometry g, const T* __restrict__ y,
                                                          const double* __restrict__ G0,
                                                          const double* __restrict__ G0b,
                                                          double* __restrict__ part) {
    __shared__ double sh_s[kAnaThreads / kWave][kNAna];
    __shared__ double sh_e[kAnaThreads / kWave][kNAna];
    double S[kNAna], E[kNAna];
#pragma unroll
    for (int q = 0; q < kNAna; ++q) { S[q] = 0.0; E[q] = 0.0; }
    const long long pts = (long long)g.nx * g.nz * g.nv;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < pts; n += stride) {
        const long long site = n / g.nv;
        const int v = (int)(n % g.nv);
        const long long o = site * g.nvp + v;
        double f[kNVar];
        bool finite = true;
#pragma unroll
        for (int q = 0; q < kNVar; ++q) {
            f[q] = (double)y[q * g.fs + o];
            finite = finite && __builtin_isfinite(f[q]);
        }
        const double w = g.vw[v], g0 = G0[v], g0b = G0b[v];
        const double iG = 1.0 / g0, iGb = 1.0 / g0b;
        const double P1 = 2.0 * f[2] * iG, P2 = -2.0 * f[3] * iG, P3 = (f[0] - f[1]) * iG;
        const double P1b = 2.0 * f[6] * iGb, P2b = 2.0 * f[7] * iGb, P3b = (f[4] - f[5]) * iGb;
        const double dP = fabs(1.0 - sqrt(P1 * P1 + P2 * P2 + P3 * P3));
        const double dPb = fabs(1.0 - sqrt(P1b * P1b + P2b * P2b + P3b * P3b));
        double term[kNAna - 2];
        term[ANA_SURV] = w * f[0];
        term[ANA_SURVB] = w * f[4];
        term[ANA_AVGP] = w * g0 * dP;
        term[ANA_AVGPB] = w * g0b * dPb;
        term[ANA_AM01] = w * (f[2] - f[6]);
        term[ANA_AM02] = w * (-f[3] - f[7]);
        term[ANA_AM03] = w * 0.5 * (f[0] - f[1] - f[4] + f[5]);
        term[ANA_NOR] = w * g0;
        term[ANA_NORB] = w * g0b;
        term[ANA_REX2] = w * (f[2] * f[2] + f[3] * f[3]);
#pragma unroll
        for (int q = 0; q < kNAna - 2; ++q) {
            double s2, err;
            two_sum(S[q], term[q], s2, err);
            S[q] = s2;
            E[q] += err;
        }
        // std::max semantics of nuosc_ana.cpp:64 (a NaN never replaces the max)
        const double m = (dP < dPb) ? dPb : dP;
        S[ANA_MAXDP] = (S[ANA_MAXDP] < m) ? m : S[ANA_MAXDP];
        if (!finite) S[ANA_BAD] += 1.0;
    }
    // wave then block reduction in a fixed order
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kNAna; ++q) {
        for (int m = 32; m >= 1; m >>= 1) {
            const double os = __shfl_xor(S[q], m), oe = __shfl_xor(E[q], m);
            if (q == ANA_MAXDP) {
                S[q] = (S[q] < os) ? os : S[q];
            } else {
                double s2, err;
                two_sum(S[q], os, s2, err);
                S[q] = s2;
                E[q] = E[q] + oe + err;
            }
        }
        if (lane == 0) { sh_s[wv][q] = S[q]; sh_e[wv][q] = E[q]; }
    }
    __syncthreads();
    if (threadIdx.x < kNAna) {
        const int q = threadIdx.x;
        double s = sh_s[0][q], e = sh_e[0][q];
        for (int k = 1; k < kAnaThreads / kWave; ++k) {
            if (q == ANA_MAXDP) {
                s = (s < sh_s[k][q]) ? sh_s[k][q] : s;
            } else {
                double s2, err;
                two_sum(s, sh_s[k][q], s2, err);
                s = s2;
                e = e + sh_e[k][q] + err;
            }
        }
        part[(long long)blockIdx.x * kNAna * 2 + q] = s;
        part[(long long)blockIdx.x * kNAna * 2 + kNAna + q] = e;
    }
}

__global__ void k_analysis_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
    const int q = threadIdx.x;
    if (q >= kNAna) return;
    double s = part[q], e = part[kNAna + q];
    for (int b = 1; b < nb; ++b) {
        const double ps = part[(long long)b * kNAna * 2 + q], pe = part[(long long)b * kNAna * 2 + kNAna + q];
        if (q == ANA_MAXDP) {
            s = (s < ps) ? ps : s;
        } else {
            double s2, err;
            two_sum(s, ps, s2, err);
            s = s2;
            e = e + pe + err;
        }
    }
    out[q] = (q == ANA_MAXDP) ? s : s + e;
}

template <typename T>
hipError_t launch_analysis(const Geometry& g, const T* y, const double* G0, const double* G0b,
                           double* part, double* out, hipStream_t s) {
    const int nb = ana_blocks(g);
    hipLaunchKernelGGL((k_analysis<T>), dim3(nb), dim3(kAnaThreads), 0, s, g, y, G0, G0b, part);
    hipLaunchKernelGGL(k_analysis_final, dim3(1), dim3(64), 0, s, part, nb, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// layout conversions, snapshots, renormalisation
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_ghosted(const Geometry g, int gx, int gz, const T* __restrict__ fld,
                          double* __restrict__ dst) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nzg = g.nz + 2 * gz;
    const long long total = (long long)(g.nx + 2 * gx) * nzg * g.nv;
    if (n >= total) return;
    const int v = (int)(n % g.nv);
    const long long s = n / g.nv;
    const int jg = (int)(s % nzg), ig = (int)(s / nzg);
    const int i = ig - gx, j = jg - gz;
    const bool xin = (i >= 0 && i < g.nx), zin = (j >= 0 && j < g.nz);
    if (!xin && !zin) { dst[n] = 0.0; return; }  // corners are never filled
    const int iw = wrapi(i, g.nx), jw = wrapi(j, g.nz);
    dst[n] = (double)fld[((long long)iw * g.nz + jw) * g.nvp + v];
}

template <typename T>
hipError_t launch_ghosted(const Geometry& g, int gx, int gz, const T* field, double* dst, hipStream_t s) {
    const long long total = (long long)(g.nx + 2 * gx) * (g.nz + 2 * gz) * g.nv;
    hipLaunchKernelGGL((k_ghosted<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, gx, gz,
                       field, dst);
    return hipGetLastError();
}

template <typename T>
__global__ void k_pack(const Geometry g, const T* __restrict__ fld, double* __restrict__ dst) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.nz * g.nv;
    if (n >= total) return;
    const long long s = n / g.nv;
    const int v = (int)(n % g.nv);
    dst[n] = (double)fld[s * g.nvp + v];
}

template <typename T>
__global__ void k_unpack(const Geometry g, const double* __restrict__ src, T* __restrict__ fld) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.nz * g.nvp;
    if (n >= total) return;
    const long long s = n / g.nvp;
    const int v = (int)(n % g.nvp);
    fld[n] = (v < g.nv) ? (T)src[s * g.nv + v] : (T)0;
}

template <typename T>
hipError_t launch_pack(const Geometry& g, const T* field, double* dst, hipStream_t s) {
    const long long total = (long long)g.nx * g.nz * g.nv;
    hipLaunchKernelGGL((k_pack<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, field, dst);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_unpack(const Geometry& g, const double* src, T* field, hipStream_t s) {
    const long long total = (long long)g.nx * g.nz * g.nvp;
    hipLaunchKernelGGL((k_unpack<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, src, field);
    return hipGetLastError();
}

template <typename T>
__global__ void k_p3_slice(const Geometry g, const T* __restrict__ y, const double* __restrict__ G0,
                           const int* __restrict__ xi, int sx, const int* __restrict__ vi, int sv,
                           double* __restrict__ dst) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)sx * g.nz * sv;
    if (n >= total) return;
    const int b = (int)(n % sv);
    const long long r = n / sv;
    const int j = (int)(r % g.nz), a = (int)(r / g.nz);
    const int v = vi[b];
    const long long o = ((long long)xi[a] * g.nz + j) * g.nvp + v;
    const double iG = 1.0 / G0[v];  // nuosc_ana.cpp:10,14
    dst[n] = ((double)y[0 * g.fs + o] - (double)y[1 * g.fs + o]) * iG;
}

template <typename T>
hipError_t launch_p3_slice(const Geometry& g, const T* y, const double* G0, const int* xi, int sx,
                           const int* vi, int sv, double* dst, hipStream_t s) {
    const long long total = (long long)sx * g.nz * sv;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL((k_p3_slice<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, y, G0,
                       xi, sx, vi, sv, dst);
    return hipGetLastError();
}

// nuosc_ana.cpp:140-166, as written there
template <typename T>
__global__ void k_renorm(const Geometry g, T* __restrict__ y, const double* __restrict__ G0,
                         const double* __restrict__ G0b) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.nz * g.nvp;
    if (n >= total) return;
    const int v = (int)(n % g.nvp);
    if (v >= g.nv) return;
    double f[kNVar];
#pragma unroll
    for (int q = 0; q < kNVar; ++q) f[q] = (double)y[q * g.fs + n];
    const double iG = 1.0 / G0[v], iGb = 1.0 / G0b[v];
    const double P1 = f[2] * iG, P2 = -f[3] * iG, P3 = (f[0] - f[1]) * iG;
    const double P1b = f[6] * iGb, P2b = f[7] * iGb, P3b = (f[4] - f[5]) * iGb;
    const double iP = 1.0 / sqrt(P1 * P1 + P2 * P2 + P3 * P3);
    const double iPb = 1.0 / sqrt(P1b * P1b + P2b * P2b + P3b * P3b);
    const double tmp = iP * P3 * G0[v], tmpb = iPb * P3b * G0b[v];
    y[0 * g.fs + n] = (T)(G0[v] + tmp);
    y[1 * g.fs + n] = (T)(G0[v] - tmp);
    y[2 * g.fs + n] = (T)(f[2] * iP);
    y[3 * g.fs + n] = (T)(f[3] * iP);
    y[4 * g.fs + n] = (T)(G0b[v] + tmpb);
    y[5 * g.fs + n] = (T)(G0b[v] - tmpb);
    y[6 * g.fs + n] = (T)(f[6] * iPb);
    y[7 * g.fs + n] = (T)(f[7] * iPb);
}

template <typename T>
hipError_t launch_renorm(const Geometry& g, T* y, const double* G0, const double* G0b, hipStream_t s) {
    const long long total = (long long)g.nx * g.nz * g.nvp;
    hipLaunchKernelGGL((k_renorm<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, y, G0, G0b);
    return hipGetLastError();
}

// explicit instantiations
#define NUOSC_INST(T)                                                                                   \
    template hipError_t launch_stage<T>(int, const Geometry&, const Physics&, const StagePlan&, const T*, \
                                        const T*, T*, T*, const double*, double*, double, double,        \
                                        hipStream_t);                                                    \
    template hipError_t launch_moments<T>(const Geometry&, const T*, double*, hipStream_t);              \
    template hipError_t launch_init<T>(const Geometry&, int, const double*, const double*, const double*, \
                                       const double*, double, uint64_t, T*, hipStream_t);                \
    template hipError_t launch_analysis<T>(const Geometry&, const T*, const double*, const double*,      \
                                           double*, double*, hipStream_t);                               \
    template hipError_t launch_ghosted<T>(const Geometry&, int, int, const T*, double*, hipStream_t);    \
    template hipError_t launch_pack<T>(const Geometry&, const T*, double*, hipStream_t);                 \
    template hipError_t launch_unpack<T>(const Geometry&, const double*, T*, hipStream_t);               \
    template hipError_t launch_p3_slice<T>(const Geometry&, const T*, const double*, const int*, int,    \
                                           const int*, int, double*, hipStream_t);                       \
    template hipError_t launch_renorm<T>(const Geometry&, T*, const double*, const double*, hipStream_t);
NUOSC_INST(double)
NUOSC_INST(float)
#undef NUOSC_INST

}  // namespace nuosc
