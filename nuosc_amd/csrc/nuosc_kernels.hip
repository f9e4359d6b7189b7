// nuosc_kernels.hip -- CDNA4 (gfx950) kernels of the nuosc RK4 time stepper.
//
// The hot path is k_stage: one launch per RK4 stage.  Each wavefront owns R
// adjacent x-rows, one chunk of 64/R directions and one z-segment, and
// marches along z keeping a 5-point z-window of all 8 fields in registers
// (every value of its rows is loaded once); the x-neighbours come from a
// small LDS slab filled with the wave's own rows and 4 halo rows, which the
// neighbouring waves of the block load as their own rows at the same time
// (L2 hits).  The state is stored site-interleaved ([x][z][field][v]) with
// periodic ghost x rows, so the rows a wave touches are one contiguous range
// under one buffer descriptor and a z-march walks through few memory pages.
// The velocity moments needed by the interaction term (a reduction over ALL
// directions of a site) are produced by the previous stage's epilogue as
// per-(site, chunk) partial sums, so no separate pass over the state is
// needed; the consumer adds the partials of a site in chunk order.
//
// RK4 is streamed with four state copies u, A, yA, yB in 12 passes over the
// 8 fields per step.  With the stage inputs Y2 = u + dt/2 k1,
// Y3 = u + dt/2 k2, Y4 = u + dt k3, the new state is
//   u + dt/6 (k1 + 2k2 + 2k3 + k4) = (A + Y4) / 3 + dt/6 k4,
//   A = Y2 + 2 Y3 - u = Y2 + u + dt k2,
// so no running sum of the k's has to be read and rewritten every stage:
//   S1: read u            -> yA = Y2
//   S2: read yA, u        -> yB = Y3,  A = Y2 + u + dt k2
//   S3: read yB, u        -> yA = Y4
//   S4: read yA, A        -> u  = (A + Y4) / 3 + dt/6 k4
// which equals the reference's combination u + dt/6 (k4 + (k1 + 2k2) + 2k3)
// (nuosc_class.cpp:374-427) up to rounding (A ~ 2u and Y4 ~ u, so no
// cancellation).
//
// The right-hand side follows nuosc_class.cpp:11-307 (2 flavours, 3rd-order
// Kreiss-Oliger, 4th-order central advection, vacuum term), with the
// reference's expression order.
#include <hip/hip_runtime.h>

#include "nuosc_dev.h"
#include "nuosc_lanes.h"

#include <algorithm>
#include <cstdint>
#include <vector>

namespace nuosc {

// Stage-kernel wave shape per precision: R x rows of 64/R lanes, V adjacent
// directions per lane.  Both precisions 1 x 64 lanes x 1: a row segment of
// one field is four 128-byte lines in FP64, two in FP32.  Wide direction
// chunks mean fewer partial moment records per site (11 at 642 directions,
// instead of 41 with 4 x 16 lanes in FP64 or 21 with 2 x 32 in FP32) and
// fewer reads of each site's moment totals, for more block-edge rows (read
// mostly from L2): at 512 x 512 x 642 the FP64 step is 6% faster with 1 x 64
// lanes than with 4 x 16 (2 x 32: 3.5%), the FP32 step 2% faster than with
// 2 x 32.  Two directions per lane (FP32) need ~250 VGPRs and ran 7% slower.
template <typename T>
struct StageShape;
// W waves per block: a block covers W R adjacent x rows of one chunk and
// loads the rows beyond its edges as halos.  4 waves in both precisions:
// blocks of 8 waves read fewer edge rows but measured slower (the per-step
// barrier waits for more waves).
template <>
struct StageShape<double> {
    static constexpr int R = 1, V = 1, W = 4;
};
template <>
struct StageShape<float> {
    static constexpr int R = 1, V = 1, W = 4;
};
template <typename T>
constexpr int kVals = StageShape<T>::V;

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Orders the LDS accesses of a wavefront: everything before is visible to
// every lane of the wave after (the wave executes in lockstep; this keeps the
// compiler from moving LDS accesses across it).
// The fences cover LDS only, so outstanding global loads (prefetches for
// later z steps) are not waited for here.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Knuth/Dekker two-sum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
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

// ---------------------------------------------------------------------------
// buffer addressing: one buffer resource per field (SGPR base), the uniform
// site offset in soffset and the lane's direction offset in voffset, so no
// per-load address arithmetic lives in VGPRs
// ---------------------------------------------------------------------------
typedef __attribute__((ext_vector_type(2))) unsigned int u32x2;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw-buffer descriptor word 3

template <typename T>
struct Lane;
template <>
struct Lane<double> {
    __device__ static __forceinline__ double ld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    }
    __device__ static __forceinline__ void st(double x, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, vo, so, 0);
    }
};
template <>
struct Lane<float> {
    __device__ static __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
    }
    __device__ static __forceinline__ void st(float x, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, vo, so, 0);
    }
};

// The stored x rows of one state copy that one wavefront reads or writes:
// rows xs0 .. xs0 + nrow - 1 under one buffer descriptor (fewer than 4 GiB).
// Offsets: lane's x row and direction in voffset, z row and field in the
// scalar offset.  The descriptor is rebuilt from the 64-bit base at each use
// (a few scalar instructions): kept live across the z loop, the descriptors
// of four state copies crowd the scalar register file.
template <typename T>
struct Rows {
    unsigned long long base;   // address of stored x row xs0
    unsigned bytes;            // nrow stored x rows
    unsigned rowb;             // bytes of one field row (nvp directions)
    // base address with an empty asm barrier: the compiler cannot hoist the
    // descriptors built from it out of the loop
    __device__ __forceinline__ unsigned long long fresh_base() const {
        unsigned long long b = base;
        asm volatile("" : "+s"(b));
        return b;
    }
    __device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(unsigned long long b) const {
        return __builtin_amdgcn_make_buffer_rsrc((void*)b, 0, (int)bytes, kRsrcWord3);
    }
    __device__ __forceinline__ unsigned foff(int f) const { return (unsigned)f * rowb; }
    // kVals<T> adjacent values per lane.  Two floats are split off their
    // 8-byte load through a 64-bit integer: taking the elements of the loaded
    // vector directly makes this compiler narrow the load to its first dword
    // and return that dword for both.  AUX = 2 marks a streaming access (the
    // nt cache policy): data read or written once per stage.
    template <int AUX = 0>
    __device__ __forceinline__ void ldv(unsigned long long b, int f, unsigned vo, unsigned so,
                                        T (&x)[kVals<T>]) const {
        if constexpr (sizeof(T) * kVals<T> == 4) {
            x[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc(b), vo, so + foff(f), AUX));
        } else if constexpr (sizeof(T) == 8) {
            x[0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc(b), vo, so + foff(f), AUX));
        } else {
            const u32x2 r = __builtin_amdgcn_raw_buffer_load_b64(rsrc(b), vo, so + foff(f), AUX);
            const unsigned long long w = __builtin_bit_cast(unsigned long long, r);
            x[0] = __builtin_bit_cast(float, (unsigned)w);
            x[1] = __builtin_bit_cast(float, (unsigned)(w >> 32));
        }
    }
    // stores stream (nt): the next reader is the next stage, a full pass later
    __device__ __forceinline__ void stv(unsigned long long b, int f, const T (&x)[kVals<T>], unsigned vo,
                                        unsigned so) const {
        if constexpr (sizeof(T) * kVals<T> == 4) {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x[0]), rsrc(b), vo, so + foff(f), 2);
        } else {
            u32x2 r;
            if constexpr (sizeof(T) == 8) {
                r = __builtin_bit_cast(u32x2, x[0]);
            } else {
                r[0] = __builtin_bit_cast(unsigned, x[0]);
                r[1] = __builtin_bit_cast(unsigned, x[1]);
            }
            __builtin_amdgcn_raw_buffer_store_b64(r, rsrc(b), vo, so + foff(f), 2);
        }
    }
};

// Value of lane n (n < 16) of x's 16-lane DPP row, in every lane of the row
// (v_mov_b32_dpp row_newbcast:n on each half).
template <int N>
__device__ __forceinline__ double row_bcast(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x150 + N, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x150 + N, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// Projection m0 - m1 vx - m2 vy - m3 vz of moments K..K+3 of a site onto a
// direction (INP, nuosc_class.cpp:187); the site's moments are held one per
// lane (lane n of each DPP row holds moment n).
template <int K>
__device__ __forceinline__ double row_project(double x, double vx, double vy, double vz) {
    return row_bcast<K>(x) - row_bcast<K + 1>(x) * vx - row_bcast<K + 2>(x) * vy - row_bcast<K + 3>(x) * vz;
}

// constant address space: uniform loads from it become scalar loads
typedef const __attribute__((address_space(4))) double cdouble;

// ---------------------------------------------------------------------------
// RK4 stage kernel
//
// A wavefront covers R adjacent x-rows x L = 64/R directions (R = 1 in both
// precisions: a row segment of one field is 512 bytes in FP64, 256 in FP32)
// and marches along z over one segment.  Per lane it keeps a 5-deep z-window of its own row for
// all 8 fields in registers.  The four waves of a block cover 4R adjacent
// rows of one direction chunk and move along z together (one LDS barrier
// per step): each writes its own rows into a block-shared two-level LDS slab
// through which all x-neighbours are read, so only the 4 rows beyond the
// block's edges are loaded as halo rows (one per wave, at z+2, staged into
// LDS one step later).  Moments of z+1 are prefetched during step z.
// ---------------------------------------------------------------------------
template <typename T>
struct StageArgs {
    Geometry g;
    unsigned long long yin, u, w, yout;   // state copies (address of stored x row 0)
    unsigned long long mtot;  // moment totals of yin: [site][kMomStride] doubles
    unsigned mtot_bytes;
    unsigned long long mpart; // partial moments of yout: [site][nch][kMomRec] (+ one spare x row)
    unsigned mpart_bytes;
    double cw, cy;
    double ca, cb;            // SSP-RK3 combination a*u + b*(y + dt*k)
    double mu2, mu, p2st, p2ct, pmo, ct2, st;
    int zlen, nquads;
};

// Sums of 16 values (with rounding errors ae carried in the compensated form)
// over each group of L consecutive lanes (L = 16, 32, 64):
// four halving exchanges (each lane keeps the half of the values selected by
// one lane-index bit) and plain butterflies for the remaining bits.  Lane d of
// a group ends with the group total of value d / (L/16).  Each level has its
// own arrays so that every select has constant indices.
template <int M>
__device__ __forceinline__ void halve(const double* in, double* out, int n2, bool hb) {
#pragma unroll
    for (int k = 0; k < n2; ++k) {
        const double lo = in[k], hi = in[k + n2];
        out[k] = (hb ? hi : lo) + __shfl_xor(hb ? lo : hi, M);
    }
}
template <int M>
__device__ __forceinline__ void halve_c(const double* in, const double* ine, double* out, double* oute, int n2,
                                        bool hb) {
#pragma unroll
    for (int k = 0; k < n2; ++k) {
        const double lo = in[k], hi = in[k + n2], loe = ine[k], hie = ine[k + n2];
        double s2, err;
        two_sum(hb ? hi : lo, __shfl_xor(hb ? lo : hi, M), s2, err);
        out[k] = s2;
        oute[k] = (hb ? hie : loe) + __shfl_xor(hb ? loe : hie, M) + err;
    }
}

template <int L, bool COMP>
__device__ __forceinline__ double group_sum16(const double (&a)[16], const double (&ae)[16], int lane) {
    constexpr int M1 = L / 2, M2 = L / 4, M3 = L / 8, M4 = L / 16;
    if constexpr (!COMP && L == 64) {
        return wave_sum16(a);   // the same tree with permlane swaps and DPP moves
    } else if constexpr (!COMP) {
        double b[8], c[4], d[2], e[1];
        halve<M1>(a, b, 8, lane & M1);
        halve<M2>(b, c, 4, lane & M2);
        halve<M3>(c, d, 2, lane & M3);
        halve<M4>(d, e, 1, lane & M4);
        double x = e[0];
#pragma unroll
        for (int m = M4 / 2; m >= 1; m >>= 1) x += __shfl_xor(x, m);
        return x;
    } else {
        double b[8], be[8], c[4], ce[4], d[2], de[2], e[1], ee[1];
        halve_c<M1>(a, ae, b, be, 8, lane & M1);
        halve_c<M2>(b, be, c, ce, 4, lane & M2);
        halve_c<M3>(c, ce, d, de, 2, lane & M3);
        halve_c<M4>(d, de, e, ee, 1, lane & M4);
        double x = e[0], xe = ee[0];
#pragma unroll
        for (int m = M4 / 2; m >= 1; m >>= 1) {
            double s2, err;
            two_sum(x, __shfl_xor(x, m), s2, err);
            xe = xe + __shfl_xor(xe, m) + err;
            x = s2;
        }
        return x + xe;
    }
}

// x rows staged per block of W waves for a stencil of radius RAD: block rows
// -RAD .. WR+RAD-1 (slots 0 .. WR+2RAD-1) and one slot that idle lanes
// write; D = L V directions each
template <typename T, int R, int W, int RAD>
struct Slab {
    static constexpr int D = (kWave / R) * kVals<T>;
    static constexpr int NS = W * R + 2 * RAD + 1;
    T v[2][kNVar][NS * D];   // two z levels (by parity) of all 8 fields
};

// superbee limiter (nuosc_class.cpp:5-8)
template <typename T>
__device__ __forceinline__ T flimiter(T a, T b) {
    const T sa = (T)((T)0 < a) - (T)(a < (T)0), sb = (T)((T)0 < b) - (T)(b < (T)0);
    return (T)0.5 * (sa + sb) * fmax(fmin((T)2 * fabs(a), fabs(b)), fmin(fabs(a), (T)2 * fabs(b)));
}

// Advection term along one axis for the non-default schemes
// (nuosc_class.cpp:127-178); m[k] / p[k] = the value k cells below / above,
// f = the scheme's factor (set up per direction), s = sign of the velocity.
// ZAX: the z axis.  The reference's MUSCL form differs between the axes in
// one sign (nuosc_class.cpp:150-157): for a negative velocity the x term ends
// in + flimiter(x0 - x-1, x1 - x0) / 2 and the z term in - flimiter(...) / 2;
// both are reproduced as written.
template <typename T, bool ZAX>
__device__ __forceinline__ T adv_axis(int adv, T f, int s, const T (&m)[4], T c0, const T (&p)[4]) {
    switch (adv) {
        case kAdvCfd2:
            return f * (p[1] - m[1]);
        case kAdvUpwind2: {   // x[-2s] - 4 x[-s] + 3 x[0]
            const T a2 = s > 0 ? m[2] : p[2], a1 = s > 0 ? m[1] : p[1];
            return f * (a2 - (T)4 * a1 + (T)3 * c0);
        }
        case kAdvLopsided: {  // -x[-3s] + 6 x[-2s] - 18 x[-s] + 10 x[0] + 3 x[s]
            const T a3 = s > 0 ? m[3] : p[3], a2 = s > 0 ? m[2] : p[2], a1 = s > 0 ? m[1] : p[1];
            const T b1 = s > 0 ? p[1] : m[1];
            return f * (-a3 + (T)6 * a2 - (T)18 * a1 + (T)10 * c0 + (T)3 * b1);
        }
        case kAdvMuscl: {     // f = -v/h
            const T up = (T)(s > 0) * (c0 + (T)0.5 * flimiter(c0 - m[1], p[1] - c0) - m[1] -
                                       (T)0.5 * flimiter(m[1] - m[2], c0 - m[1]));
            const T lc = (T)0.5 * flimiter(c0 - m[1], p[1] - c0);
            const T dn = (T)(s < 0) * (p[1] - (T)0.5 * flimiter(p[1] - c0, p[2] - p[1]) - c0 + (ZAX ? -lc : lc));
            return f * (up + dn);
        }
        case kAdvOff:
            return (T)0;
        default:              // 4th-order central
            return f * ((m[2] - p[2]) - (T)8.0 * (m[1] - p[1]));
    }
}

// Orders the LDS accesses of a block: every LDS write of every wave before it
// is visible to every LDS read after it.  The fences cover LDS only, so
// outstanding global loads and stores are not waited for.
__device__ __forceinline__ void block_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One wavefront's work.  Block lblock = (group of W x-row quads, direction
// chunk, z segment), quad group fastest; wave wib of the block takes quad
// W * group + wib.  Lane (g, dl) of a wave handles x row i0 + g and the V
// adjacent directions ch * D + dl * V + e, e < V.
// NOX (grids of one x row): the x neighbours of every point are its own
// value (the periodic ghost rows are copies of row 0), so no LDS slab, edge
// rows or barrier are needed, and the block's W R row groups work on W R
// consecutive direction chunks of row 0 instead of W R rows of one chunk.
// The x terms are still evaluated, from the centre value, as the reference
// does with its ghost copies.
template <typename T, int R, int W, int RAD, int STAGE, bool COMP, bool NOX = false>
__device__ __forceinline__ void stage_body(const StageArgs<T>& a, int lblock, int wib, Slab<T, R, W, RAD>& slab) {
    constexpr int NZW = 2 * RAD + 1;   // z window
    constexpr int V = kVals<T>;    // directions per lane (one 8-byte load)
    constexpr int L = kWave / R;   // lanes per row
    constexpr int D = L * V;       // directions per row chunk
    auto& xs = slab.v;

    const int lane = threadIdx.x & 63;
    const int nx = a.g.nx, nz = a.g.nz, nch = a.g.nch, nv = a.g.nv;
    const int nzs = a.g.nzs, zoff = a.g.zoff;             // stored rows per x row, ghost rows below z = 0
    const unsigned rowb = (unsigned)a.g.nvp * sizeof(T);  // bytes per field row
    const unsigned siteb = (unsigned)(a.g.srec * (long long)sizeof(T));   // site-record stride in bytes
    const unsigned xrowb = (unsigned)(a.g.xrow * (long long)sizeof(T));   // x-row stride in bytes
    // The four waves of a block sweep 4R adjacent x rows of one chunk side by
    // side, so most halo rows a wave reads are rows its neighbours in the
    // block read at about the same time (L2 hits).  Quad group fastest across
    // blocks: the neighbouring blocks, which the XCD remap keeps on one XCD
    // and starts at about the same time, own the block's edge halo rows.
    // (Chunk fastest, which keeps the chunks of a site record together,
    // measured 3-7% slower.)
    const int nqg = (a.nquads + W - 1) / W;
    const int ib = (lblock % nqg) * W * R;   // first x row of the block
    const int rest = lblock / nqg;
    const int nchb = NOX ? (nch + W * R - 1) / (W * R) : nch;   // chunk groups of the block grid
    const int chb = rest % nchb;
    const int zs = rest / nchb;
    // (a wave whose quad lies past the last one still runs, with all its
    // rows past nx: every wave of the block takes part in each step's barrier)
    const int i0 = NOX ? 0 : ib + wib * R;
    const int g = lane / L, dl = lane % L;
    const int ch = NOX ? chb * W * R + wib * R + g : chb;   // the lane's direction chunk
    const int p = ch * L + dl;           // direction group: directions p*V .. p*V+V-1
    const int r = NOX ? 0 : i0 + g;
    const bool row_ok = r < nx;
    const bool act = row_ok && p * V < nv;   // nv is a multiple of V
    const int pl = min(p, nv / V - 1);       // loads stay inside the directions

    // Block row b (x row ib + b) is stored row ib + b + kXg (the periodic
    // ghost rows make the block's rows ib-RAD .. ib+WR+RAD-1 contiguous);
    // rows past the last interior row + kXg map to the spare stored row
    // nx + 2kXg.  One descriptor per state copy covers stored rows
    // ib .. min(ib + WR + 2kXg, nxs - 1); a lane's stored row relative to ib
    // goes into its vector offset.
    const int spare = nx + 2 * kXg - ib;     // spare stored row, relative to ib
    const unsigned long long woff = (unsigned long long)ib * xrowb;
    const unsigned wbytes = (unsigned)(min(W * R + 2 * kXg + 1, a.g.nxs - ib)) * xrowb;
    const Rows<T> yin{a.yin + woff, wbytes, rowb}, uu{a.u + woff, wbytes, rowb};
    const Rows<T> ww{a.w + woff, wbytes, rowb}, yout{a.yout + woff, wbytes, rowb};
    // stored z row of interior z index jj (periodic: wrapped; open: the ghost
    // rows hold jj in [-zoff, nz + zoff))
    auto zrow = [&](int jj) { return (unsigned)wrapi(jj + zoff, nzs); };
    const unsigned dof = (unsigned)(pl * V) * sizeof(T);
    const int own = NOX ? 0 : wib * R + g;   // block row of the lane
    const unsigned vo = (unsigned)min(own + kXg, spare) * xrowb + dof;
    // Stores are unconditional (no branch, so the compiler can count the
    // outstanding memory operations across iterations exactly): a lane past
    // the last x row writes the spare row, which no lane that stores a result
    // reads; a lane past nv works on the last direction (pl) throughout and
    // stores the same bits to the same address as that direction's lane, so
    // the padding columns are never written.
    // (NOX: lanes of chunks past the last one sit in other waves than the
    // owner of direction nv-1 and store into the spare row: the SSP-RK3
    // stage 7 rewrites u in place, so a duplicate store could reach u before
    // the owner has read it)
    const unsigned vos = (unsigned)(row_ok && (!NOX || ch < nch) ? own + kXg : spare) * xrowb + dof;
    // Edge rows: edge e = wib + W g (row group g < LG of wave wib, e < 2 RAD)
    // loads block row e - RAD (e < RAD) or WR + e - RAD with its L lanes.  The
    // other lanes' offset lies past the descriptor's range, so they fetch
    // nothing (the load returns zeros), and they write the idle slot.  With
    // one row group per wave and more edge rows than waves (radius 3, one x
    // row per wave), the loaders of edges W LG .. 2 RAD - 1 load a second one.
    constexpr int NE = 2 * RAD;
    constexpr int LG = R >= 2 ? 2 : 1;
    constexpr bool E2 = NE > W * LG;
    auto edge_row = [&](int e, bool ok, unsigned& voff, int& slot) {
        const int eb = e < RAD ? e - RAD : W * R + e - RAD;   // block row
        const int erow = min(eb + kXg, spare);                  // stored row relative to ib
        voff = ok ? (unsigned)erow * xrowb + dof : wbytes;
        slot = (ok ? eb + RAD : W * R + 2 * RAD) * D + dl * V;
    };
    unsigned vh, vh2;
    int slot_h, slot_h2;
    edge_row(wib + W * g, g < LG && wib + W * g < NE, vh, slot_h);
    edge_row(W * LG + wib + W * g, E2 && g < LG && W * LG + wib + W * g < NE, vh2, slot_h2);
    const int slot_c = (own + RAD) * D + dl * V;

    // per-direction constants (nuosc_class.cpp:119-121, 180-181)
    double vxd[V], vyd[V], vzd[V], q[V][4];
    T fx[V], fz[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int ve = pl * V + e;
        vxd[e] = a.g.vx[ve];
        vyd[e] = a.g.vy[ve];
        vzd[e] = a.g.vz[ve];
        const double vwd = act ? a.g.vw[ve] : 0.0;
        if constexpr (RAD == 2) {   // nuosc_class.cpp:180-181
            fx[e] = (T)(-vxd[e] / a.g.twelve_dx);
            fz[e] = (T)(-vzd[e] / a.g.twelve_dz);
        } else {                    // the scheme's factor (nuosc_class.cpp:150-181)
            const int sx = (0.0 < vxd[e]) - (vxd[e] < 0.0), sz = (0.0 < vzd[e]) - (vzd[e] < 0.0);
            switch (a.g.adv) {
                case kAdvCfd2: fx[e] = (T)(-vxd[e] / (2 * a.g.dx)); fz[e] = (T)(-vzd[e] / (2 * a.g.dz)); break;
                case kAdvUpwind2:
                    fx[e] = (T)(-sx * vxd[e] / (2 * a.g.dx));
                    fz[e] = (T)(-sz * vzd[e] / (2 * a.g.dz));
                    break;
                case kAdvLopsided:
                    fx[e] = (T)(-sx * vxd[e] / (12 * a.g.dx));
                    fz[e] = (T)(-sz * vzd[e] / (12 * a.g.dz));
                    break;
                case kAdvMuscl: fx[e] = (T)(-vxd[e] / a.g.dx); fz[e] = (T)(-vzd[e] / a.g.dz); break;
                default: fx[e] = (T)(-vxd[e] / a.g.twelve_dx); fz[e] = (T)(-vzd[e] / a.g.twelve_dz); break;
            }
        }
        q[e][0] = vwd;
        q[e][1] = vxd[e] * vwd;
        q[e][2] = vyd[e] * vwd;
        q[e][3] = vzd[e] * vwd;
    }
    // Kreiss-Oliger: order 3 -ko/(16h), order 5 -ko/(64h) (nuosc_class.cpp:114-121)
    const T kx = (T)(RAD == 3 && a.g.ko5 ? 0.25 * a.g.ko_x : a.g.ko_x);
    const T kz = (T)(RAD == 3 && a.g.ko5 ? 0.25 * a.g.ko_z : a.g.ko_z);
    const int sxv = (0.0 < vxd[0]) - (vxd[0] < 0.0), szv = (0.0 < vzd[0]) - (vzd[0] < 0.0);
    const T mu2 = (T)a.mu2, mu = (T)a.mu, p2st = (T)a.p2st, p2ct = (T)a.p2ct;
    const T pmo = (T)a.pmo, ct2 = (T)a.ct2, st = (T)a.st;
    const T cw = (T)a.cw, cy = (T)a.cy, ca = (T)a.ca, cb = (T)a.cb;
    const T third = (T)(1.0 / 3.0);
    // moment totals of the lane's site (a lane past the last row reads row
    // nx-1): lane dl holds value dl % 16 of its row's record; every data row
    // (16 or 32 lanes) is made of whole 16-lane DPP rows, so row_project()
    // reaches all 12 values of the lane's site
    const unsigned mbase = (unsigned)ib * nz;             // record of (ib, z = 0)
    const unsigned vom = (unsigned)(min(r, nx - 1) - ib) * (unsigned)nz * (kMomStride * 8) + (unsigned)(dl & 15) * 8;

    const int j0 = zs * a.zlen;
    const int j1 = min(nz, j0 + a.zlen);

    T win[kNVar][NZW][V]; // own row, z = j-RAD .. j+RAD
    T hreg[kNVar][V];     // edge row of z = j+2, staged into LDS in step j+1
    T hreg2[E2 ? kNVar : 1][V];   // the second edge row (E2)
    // velocity moments of the stage input at this lane's site (one value per
    // lane) for z = j, loaded at the end of step j-1
    unsigned long long bm = a.mtot;
    asm volatile("" : "+s"(bm));
    const __amdgpu_buffer_rsrc_t rm = __builtin_amdgcn_make_buffer_rsrc((void*)bm, 0, (int)a.mtot_bytes, kRsrcWord3);
    auto load_moment = [&](int jj) {
        const unsigned sm = (mbase + jj) * (kMomStride * 8);
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rm, vom, sm, 0));
    };
#pragma unroll
    for (int qz = 0; qz < 2 * RAD; ++qz) {
        const unsigned so = zrow(j0 - RAD + qz) * siteb;
#pragma unroll
        for (int f = 0; f < kNVar; ++f) yin.ldv(yin.base, f, vo, so, win[f][qz]);
    }
    // edge row of z = j0 straight into its level; z = j0+1 into hreg (staged
    // in the first step)
    if constexpr (!NOX) {
        const unsigned so = zrow(j0) * siteb, s1 = zrow(j0 + 1) * siteb;
#pragma unroll
        for (int f = 0; f < kNVar; ++f) {
            T x[V];
            yin.ldv(yin.base, f, vh, so, x);
#pragma unroll
            for (int e = 0; e < V; ++e) xs[j0 & 1][f][slot_h + e] = x[e];
        }
#pragma unroll
        for (int f = 0; f < kNVar; ++f) yin.ldv(yin.base, f, vh, s1, hreg[f]);
        if constexpr (E2) {
#pragma unroll
            for (int f = 0; f < kNVar; ++f) {
                T x[V];
                yin.ldv(yin.base, f, vh2, so, x);
#pragma unroll
                for (int e = 0; e < V; ++e) xs[j0 & 1][f][slot_h2 + e] = x[e];
            }
#pragma unroll
            for (int f = 0; f < kNVar; ++f) yin.ldv(yin.base, f, vh2, s1, hreg2[f]);
        }
    }
    double Ml = load_moment(j0);

    for (int j = j0; j < j1; ++j) {
        const unsigned so = (unsigned)(j + zoff) * siteb;
        const int lv = j & 1;
        if constexpr (!NOX) {
            // own rows of z = j into this level (the edge rows of z = j are there)
#pragma unroll
            for (int f = 0; f < kNVar; ++f)
#pragma unroll
                for (int e = 0; e < V; ++e) xs[lv][f][slot_c + e] = win[f][RAD][e];
            // all rows of z = j are in level lv, and every wave is done with step
            // j-1, whose level receives the edge rows of z = j+1 (loaded in step
            // j-1; no wave reads that level before the next barrier)
            block_sync();
#pragma unroll
            for (int f = 0; f < kNVar; ++f)
#pragma unroll
                for (int e = 0; e < V; ++e) xs[lv ^ 1][f][slot_h + e] = hreg[f][e];
            if constexpr (E2) {
#pragma unroll
                for (int f = 0; f < kNVar; ++f)
#pragma unroll
                    for (int e = 0; e < V; ++e) xs[lv ^ 1][f][slot_h2 + e] = hreg2[f][e];
            }
        }
        // ---- the step's loads, in the order they are consumed: own row
        // z+RAD (stencil), u and w (update), edge row z+2 (staged in the next
        // step); the moments of z+1 are loaded at the end of the step
        const unsigned long long by = yin.fresh_base();
        const unsigned s2 = zrow(j + 2) * siteb;
#pragma unroll
        for (int f = 0; f < kNVar; ++f) yin.ldv(by, f, vo, zrow(j + RAD) * siteb, win[f][2 * RAD]);
        // (scheduling barriers keep the groups in this issue order, so each
        // wait below covers only the groups needed so far)
        __builtin_amdgcn_sched_barrier(0);
        T uc[kNVar][V], wc[kNVar][V];
        if constexpr (STAGE == 2 || STAGE == 3 || STAGE == 6 || STAGE == 7) {
            const unsigned long long bu = uu.fresh_base();
#pragma unroll
            for (int f = 0; f < kNVar; ++f) uu.template ldv<2>(bu, f, vo, so, uc[f]);
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (STAGE == 4) {
            const unsigned long long bw = ww.fresh_base();
#pragma unroll
            for (int f = 0; f < kNVar; ++f) ww.template ldv<2>(bw, f, vo, so, wc[f]);
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (!NOX) {
#pragma unroll
            for (int f = 0; f < kNVar; ++f) yin.ldv(by, f, vh, s2, hreg[f]);
            if constexpr (E2) {
#pragma unroll
                for (int f = 0; f < kNVar; ++f) yin.ldv(by, f, vh2, s2, hreg2[f]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }

        T yo[kNVar][V], wn[kNVar][V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            // projections onto direction v (INP, nuosc_class.cpp:187)
            const T A = (T)row_project<0>(Ml, vxd[e], vyd[e], vzd[e]);
            const T B = (T)row_project<4>(Ml, vxd[e], vyd[e], vzd[e]);
            const T C = (T)row_project<8>(Ml, vxd[e], vyd[e], vzd[e]);

            const T ee = win[0][RAD][e], mm = win[1][RAD][e], emr = win[2][RAD][e], emi = win[3][RAD][e];
            const T bee = win[4][RAD][e], bmm = win[5][RAD][e], bemr = win[6][RAD][e], bemi = win[7][RAD][e];
            // interaction (nuosc_class.cpp:193-200)
            const T Iee = mu2 * (emi * A + emr * B);
            const T Iemr = -mu * ((ee - mm) * B + C * emi);
            const T Iemi = mu * ((mm - ee) * A + C * emr);
            const T Ibee = mu2 * (bemi * A + bemr * B);
            const T Ibemr = -mu * ((bee - bmm) * B + C * bemi);
            const T Ibemi = mu * ((bmm - bee) * A + C * bemr);
            // local terms (interaction + vacuum, nuosc_class.cpp:280-287)
            const T loc[kNVar] = {Iee, -Iee, Iemr, Iemi, Ibee, -Ibee, Ibemr, Ibemi};
            const T vac[kNVar] = {-(p2st * emi), p2st * emi, -(p2ct * emi), pmo * (ct2 * emr + st * (ee - mm)),
                                  -(p2st * bemi), p2st * bemi, -(p2ct * bemi),
                                  pmo * (ct2 * bemr + st * (bee - bmm))};
#pragma unroll
            for (int f = 0; f < kNVar; ++f) {
                const T c0 = win[f][RAD][e];
                T adv, ko;
                if constexpr (RAD == 2) {   // the reference's default discretisation
                    const T xm2 = NOX ? c0 : xs[lv][f][slot_c - 2 * D + e];
                    const T xm1 = NOX ? c0 : xs[lv][f][slot_c - D + e];
                    const T xp1 = NOX ? c0 : xs[lv][f][slot_c + D + e];
                    const T xp2 = NOX ? c0 : xs[lv][f][slot_c + 2 * D + e];
                    const T zm2 = win[f][0][e], zm1 = win[f][1][e], zp1 = win[f][3][e], zp2 = win[f][4][e];
                    // advection (nuosc_class.cpp:182-183)
                    adv = fz[e] * ((zm2 - zp2) - (T)8.0 * (zm1 - zp1)) + fx[e] * ((xm2 - xp2) - (T)8.0 * (xm1 - xp1));
                    // Kreiss-Oliger (nuosc_class.cpp:122-123)
                    ko = kz * (zm2 + zp2 - (T)4 * (zm1 + zp1) + (T)6 * c0) +
                         kx * (xm2 + xp2 - (T)4 * (xm1 + xp1) + (T)6 * c0);
                } else {                    // scheme chosen at run time (nuosc_class.cpp:112-183)
                    T xm[4], xp[4], zm[4], zp[4];
#pragma unroll
                    for (int kk = 1; kk <= RAD; ++kk) {
                        xm[kk] = NOX ? c0 : xs[lv][f][slot_c - kk * D + e];
                        xp[kk] = NOX ? c0 : xs[lv][f][slot_c + kk * D + e];
                        zm[kk] = win[f][RAD - kk][e];
                        zp[kk] = win[f][RAD + kk][e];
                    }
                    xm[0] = xp[0] = zm[0] = zp[0] = c0;
                    adv = adv_axis<T, true>(a.g.adv, fz[e], szv, zm, c0, zp) + adv_axis<T, false>(a.g.adv, fx[e], sxv, xm, c0, xp);
                    if (a.g.ko5)
                        ko = kz * (zm[3] + zp[3] - (T)6 * (zm[2] + zp[2]) + (T)15 * (zm[1] + zp[1]) - (T)20 * c0) +
                             kx * (xm[3] + xp[3] - (T)6 * (xm[2] + xp[2]) + (T)15 * (xm[1] + xp[1]) - (T)20 * c0);
                    else
                        ko = kz * (zm[2] + zp[2] - (T)4 * (zm[1] + zp[1]) + (T)6 * c0) +
                             kx * (xm[2] + xp[2] - (T)4 * (xm[1] + xp[1]) + (T)6 * c0);
                }
                const T k = loc[f] + adv + ko + vac[f];
                if constexpr (STAGE == 0) {
                    yo[f][e] = k;
                } else if constexpr (STAGE == 1) {   // Y2 = u + dt/2 k1
                    yo[f][e] = c0 + cy * k;
                } else if constexpr (STAGE == 2) {   // A = Y2 + u + dt k2, Y3 = u + dt/2 k2
                    wn[f][e] = (c0 + uc[f][e]) + cw * k;
                    yo[f][e] = uc[f][e] + cy * k;
                } else if constexpr (STAGE == 3) {   // Y4 = u + dt k3
                    yo[f][e] = uc[f][e] + cy * k;
                } else if constexpr (STAGE == 4) {   // u' = (A + Y4) / 3 + dt/6 k4
                    yo[f][e] = (wc[f][e] + c0) * third + cw * k;
                } else if constexpr (STAGE == 5) {   // SSP-RK3 (nuosc_class.cpp:436)
                    yo[f][e] = c0 + cy * k;
                } else {                             // SSP-RK3 (nuosc_class.cpp:440, 444, 369)
                    yo[f][e] = ca * uc[f][e] + cb * (c0 + cy * k);
                }
            }
        }
        {
            const unsigned long long bw = ww.fresh_base();
            const unsigned long long bo = yout.fresh_base();
#pragma unroll
            for (int f = 0; f < kNVar; ++f) {
                if constexpr (STAGE == 2) ww.stv(bw, f, wn[f], vos, so);
                yout.stv(bo, f, yo[f], vos, so);
            }
        }
        if constexpr (STAGE != 0) {
            // moment terms of the lane's V directions, summed per lane first
            // (exactly, with the rounding error carried, in the FP32 form)
            double t[16], te[16];
            {
                T y0[kNVar];
#pragma unroll
                for (int f = 0; f < kNVar; ++f) y0[f] = yo[f][0];
                moment_terms<T>(y0, q[0], t);
            }
#pragma unroll
            for (int s = 0; s < 16; ++s) te[s] = 0.0;
#pragma unroll
            for (int e = 1; e < V; ++e) {
                T ye[kNVar];
#pragma unroll
                for (int f = 0; f < kNVar; ++f) ye[f] = yo[f][e];
                double u[16];
                moment_terms<T>(ye, q[e], u);
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    if constexpr (COMP) {
                        double s2, err;
                        two_sum(t[s], u[s], s2, err);
                        t[s] = s2;
                        te[s] += err;
                    } else {
                        t[s] += u[s];
                    }
                }
            }
            // lane dl of a row holds the row total of value dl >> sh (values
            // 12..15 are zeros, stored into the spare row's records, which
            // nothing reads; lanes sharing a value store the same bits).
            // The sum over the row's lanes is an FP64 pairwise tree (log2 L
            // levels: its error is at most log2(L) 2^-53 of the summed
            // magnitudes), compensated only where a lane carries several
            // directions' rounding errors; the chunk records of a site are
            // added with compensation (k_moment_totals) in the FP32 form.
            // (A compensated tree measured 13% slower in FP32 stage 1 and
            // changes the moments by ~1e-16 relative, far below FP32
            // rounding of the state.)
            const double tot = group_sum16<L, COMP && (V > 1)>(t, te, lane);
            constexpr int sh = (L == 16) ? 0 : (L == 32 ? 1 : 2);
            unsigned long long bp = a.mpart;
            asm volatile("" : "+s"(bp));
            const __amdgpu_buffer_rsrc_t rp =
                __builtin_amdgcn_make_buffer_rsrc((void*)bp, 0, (int)a.mpart_bytes, kRsrcWord3);
            const int sv = dl >> sh;                       // value index 0..15
            const bool used = sv < kNMom;
            // (NOX: row groups past the last chunk write the spare row's records)
            const bool rec_ok = row_ok && used && (!NOX || ch < nch);
            const unsigned rec =
                ((unsigned)((rec_ok ? r : nx) - ib) * (unsigned)nz * nch + (unsigned)(NOX ? min(ch, nch - 1) : ch)) *
                (kMomRec * 8);
            const unsigned srec = (mbase + j) * (unsigned)(nch * kMomRec * 8);
            Lane<double>::st(tot, rp, rec + (unsigned)(used ? sv : sv - kNMom) * 8, srec);
        }
#pragma unroll
        for (int f = 0; f < kNVar; ++f)
#pragma unroll
            for (int e = 0; e < V; ++e)
#pragma unroll
                for (int qz = 0; qz < NZW - 1; ++qz) win[f][qz][e] = win[f][qz + 1][e];
        // moments of z+1, issued last and into the variable the next step
        // reads (so no register copy waits for it); unconditional (a branch
        // here would make every later wait a full drain): the last row of the
        // box reloads itself
        __builtin_amdgcn_sched_barrier(0);
        Ml = load_moment(min(j + 1, nz - 1));
    }
}

// Waves per SIMD the register allocation aims at: the FP64 slab kernels are
// held to 2 by their LDS (two 73.7 KB blocks of 4 waves per CU); the FP64
// one-row (NOX) stages 0-5 fit 168 VGPRs (3 waves), the SSP-RK3 stages 6 and
// 7 would spill there (2 waves); the FP32 stages fit 128 (4 waves).
template <typename T, int STAGE, int RAD = 2, bool NOX = false>
constexpr int stage_waves() {
    return RAD == 3 ? (sizeof(T) == 4 ? 2 : 1) : (sizeof(T) == 4 ? 4 : (STAGE <= 5 && NOX ? 3 : 2));
}

// RAD = 2: the reference's default discretisation; RAD = 3: any of the
// schemes, chosen at run time (Geometry::ko5, Geometry::adv)
template <typename T, int R, int STAGE, bool COMP, int RAD = 2, bool NOX = false>
__global__ __launch_bounds__(64 * StageShape<T>::W) __attribute__((amdgpu_waves_per_eu(stage_waves<T, STAGE, RAD, NOX>()))) void k_stage(const StageArgs<T> a) {
    constexpr int W = StageShape<T>::W;
    __shared__ Slab<T, R, W, RAD> slab;   // (never touched by NOX, which the compiler then leaves unallocated)
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    stage_body<T, R, W, RAD, STAGE, COMP, NOX>(a, xcd_remap(blockIdx.x, gridDim.x), wib, slab);
}

// Many independent runs in one launch: logical block -> (run, block of run)
// from bmap; each run's arguments from the runs table (read with scalar loads).
template <typename T, int R, int STAGE, bool COMP>
__global__ __launch_bounds__(64 * StageShape<T>::W) __attribute__((amdgpu_waves_per_eu(stage_waves<T, STAGE>()))) void k_stage_batch(
    const StageArgs<T>* __restrict__ runs, const int* __restrict__ bmap) {
    constexpr int W = StageShape<T>::W;
    __shared__ Slab<T, R, W, 2> slab;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const __attribute__((address_space(4))) int* cm = (const __attribute__((address_space(4))) int*)bmap;
    const int run = cm[2 * lb], blk = cm[2 * lb + 1];
    const StageArgs<T> a = runs[run];   // uniform, read-only: scalar loads
    stage_body<T, R, W, 2, STAGE, COMP>(a, blk, wib, slab);
}

// The same work as a persistent kernel: a grid of resident blocks takes
// logical blocks from a work queue until it is empty.  Logical blocks keep
// the split of xcd_remap (one contiguous range per XCD, block b of the grid
// runs on XCD b % 8), and each XCD's blocks draw from their range through
// their own counter ctr[x], so neighbouring logical blocks still run on one
// XCD at about the same time.  Every block leaves the loop once its XCD's
// counter passes the range (the exit every wave reaches); the counters are
// zeroed by k_zero_ints before each launch.
template <typename T, int R, int STAGE, bool COMP>
__global__ __launch_bounds__(64 * StageShape<T>::W) __attribute__((amdgpu_waves_per_eu(stage_waves<T, STAGE>()))) void k_stage_batch_q(
    const StageArgs<T>* __restrict__ runs, const int* __restrict__ bmap, int* __restrict__ ctr, int nitems) {
    constexpr int W = StageShape<T>::W;
    __shared__ Slab<T, R, W, 2> slab;
    __shared__ int next;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x = blockIdx.x % 8;
    const int per = nitems / 8, rem = nitems % 8;
    const int start = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    const int count = per + (x < rem ? 1 : 0);
    const __attribute__((address_space(4))) int* cm = (const __attribute__((address_space(4))) int*)bmap;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(&ctr[x], 1);
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();   // every wave has read next before thread 0 rewrites it
        if (k >= count) break;
        const int lb = start + k;
        const int run = cm[2 * lb], blk = cm[2 * lb + 1];
        const StageArgs<T> a = runs[run];   // uniform, read-only: scalar loads
        stage_body<T, R, W, 2, STAGE, COMP>(a, blk, wib, slab);
    }
}

__global__ void k_zero_ints(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int rows_per_wave(int esize) { return esize == 8 ? StageShape<double>::R : StageShape<float>::R; }
int waves_per_block(int esize) { return esize == 8 ? StageShape<double>::W : StageShape<float>::W; }
int dirs_per_chunk(int esize) {
    return esize == 8 ? (kWave / StageShape<double>::R) * StageShape<double>::V
                      : (kWave / StageShape<float>::R) * StageShape<float>::V;
}

StagePlan plan_stage(const Geometry& g, int num_cus) {
    // enough waves to give every SIMD several (>= 12 per CU), segments long
    // enough that the window warm-up stays a small overhead
    StagePlan p{};
    const int R = g.rows, W = g.wpb;
    const long long quads = (g.nx + R - 1) / R;
    // NOX: the blocks of one z segment cover the chunks in groups of W R
    const long long cgroups = g.nox ? (g.nch + W * R - 1) / (W * R) : g.nch;
    const long long rows = g.nox ? cgroups * W : quads * g.nch;
    const long long want = (long long)num_cus * 48;
    int zlen = g.nz;
    while (zlen > 8 && rows * ((g.nz + zlen - 1) / zlen) < want) zlen = (zlen + 1) / 2;
    if (zlen > 64 && rows * ((g.nz + 63) / 64) >= want) zlen = 64;
    p.zlen = zlen;
    p.nseg = (g.nz + zlen - 1) / zlen;
    p.nquads = (int)quads;
    p.nblocks = (int)(((quads + W - 1) / W) * cgroups * p.nseg);
    p.ntask = p.nblocks * W;
    return p;
}

StagePlan plan_stage_zlen(const Geometry& g, int zlen) {
    StagePlan p{};
    const int R = g.rows, W = g.wpb;
    const long long quads = (g.nx + R - 1) / R;
    const long long cgroups = g.nox ? (g.nch + W * R - 1) / (W * R) : g.nch;
    p.zlen = std::max(1, std::min(zlen, g.nz));
    p.nseg = (g.nz + p.zlen - 1) / p.zlen;
    p.nquads = (int)quads;
    p.nblocks = (int)(((quads + W - 1) / W) * cgroups * p.nseg);
    p.ntask = p.nblocks * W;
    return p;
}

bool buffer_addressable(const Geometry& g, int esize) {
    const long long lim = (1ll << 32) - 1;
    const int R = rows_per_wave(esize), W = waves_per_block(esize);
    // the block's rows plus one: an idle lane's offset (the range's end) plus
    // a scalar offset must not wrap
    const long long wave_rows = (std::min<long long>(W * R + 2 * kXg + 1, g.nxs) + 1) * g.xrow * esize;
    const long long part = (long long)(g.nx + 1) * g.nz * g.nch * kMomRec * (long long)sizeof(double);
    const long long tot = (long long)g.nx * g.nz * kMomStride * (long long)sizeof(double);
    return wave_rows < lim && part < lim && tot < lim;
}

template <typename T>
static StageArgs<T> make_stage_args(const Geometry& g, const Physics& ph, const StagePlan& pl, const T* yin,
                                    const T* u, T* w, T* yout, const double* mtot, double* mpart, double cw,
                                    double cy, double ca = 0.0, double cb = 0.0) {
    StageArgs<T> a;
    a.g = g;
    a.yin = (unsigned long long)yin;
    a.u = (unsigned long long)(u ? u : yin);
    a.w = (unsigned long long)(w ? w : yout);
    a.yout = (unsigned long long)yout;
    a.mtot = (unsigned long long)mtot;
    a.mtot_bytes = (unsigned)((long long)g.nx * g.nz * kMomStride * sizeof(double));
    a.mpart = (unsigned long long)mpart;
    a.mpart_bytes = (unsigned)((long long)(g.nx + 1) * g.nz * g.nch * kMomRec * sizeof(double));
    a.cw = cw;
    a.cy = cy;
    a.ca = ca;
    a.cb = cb;
    a.mu2 = 2 * ph.mu;
    a.mu = ph.mu;
    a.p2st = ph.pmo * 2 * ph.st;
    a.p2ct = ph.pmo * 2 * ph.ct;
    a.pmo = ph.pmo;
    a.ct2 = 2 * ph.ct;
    a.st = ph.st;
    a.zlen = pl.zlen;
    a.nquads = pl.nquads;
    return a;
}

template <typename T>
hipError_t launch_stage(int stage, const Geometry& g, const Physics& ph, const StagePlan& pl,
                        const T* yin, const T* u, T* w, T* yout, const double* mtot, double* mpart,
                        double cw, double cy, hipStream_t s, double ca, double cb) {
    if (!buffer_addressable(g, sizeof(T))) return hipErrorInvalidValue;
    const StageArgs<T> a = make_stage_args<T>(g, ph, pl, yin, u, w, yout, mtot, mpart, cw, cy, ca, cb);
    constexpr bool COMP = sizeof(T) == 4;
    constexpr int R = StageShape<T>::R;
    if (g.lpr != dirs_per_chunk((int)sizeof(T)) || g.wpb != StageShape<T>::W) return hipErrorInvalidValue;
    dim3 grid(pl.nblocks), block(64 * StageShape<T>::W);
    if (g.nox) {   // one x row, default discretisation
        switch (stage) {
            case 0: hipLaunchKernelGGL((k_stage<T, R, 0, COMP, 2, true>), grid, block, 0, s, a); break;
            case 1: hipLaunchKernelGGL((k_stage<T, R, 1, COMP, 2, true>), grid, block, 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_stage<T, R, 2, COMP, 2, true>), grid, block, 0, s, a); break;
            case 3: hipLaunchKernelGGL((k_stage<T, R, 3, COMP, 2, true>), grid, block, 0, s, a); break;
            case 4: hipLaunchKernelGGL((k_stage<T, R, 4, COMP, 2, true>), grid, block, 0, s, a); break;
            case 5: hipLaunchKernelGGL((k_stage<T, R, 5, COMP, 2, true>), grid, block, 0, s, a); break;
            case 6: hipLaunchKernelGGL((k_stage<T, R, 6, COMP, 2, true>), grid, block, 0, s, a); break;
            case 7: hipLaunchKernelGGL((k_stage<T, R, 7, COMP, 2, true>), grid, block, 0, s, a); break;
            default: return hipErrorInvalidValue;
        }
    } else if (scheme_radius(g.ko5, g.adv) == 2 && g.adv == kAdvCfd4) {
        switch (stage) {
            case 0: hipLaunchKernelGGL((k_stage<T, R, 0, COMP>), grid, block, 0, s, a); break;
            case 1: hipLaunchKernelGGL((k_stage<T, R, 1, COMP>), grid, block, 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_stage<T, R, 2, COMP>), grid, block, 0, s, a); break;
            case 3: hipLaunchKernelGGL((k_stage<T, R, 3, COMP>), grid, block, 0, s, a); break;
            case 4: hipLaunchKernelGGL((k_stage<T, R, 4, COMP>), grid, block, 0, s, a); break;
            case 5: hipLaunchKernelGGL((k_stage<T, R, 5, COMP>), grid, block, 0, s, a); break;
            case 6: hipLaunchKernelGGL((k_stage<T, R, 6, COMP>), grid, block, 0, s, a); break;
            case 7: hipLaunchKernelGGL((k_stage<T, R, 7, COMP>), grid, block, 0, s, a); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (stage) {
            case 0: hipLaunchKernelGGL((k_stage<T, R, 0, COMP, 3>), grid, block, 0, s, a); break;
            case 1: hipLaunchKernelGGL((k_stage<T, R, 1, COMP, 3>), grid, block, 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_stage<T, R, 2, COMP, 3>), grid, block, 0, s, a); break;
            case 3: hipLaunchKernelGGL((k_stage<T, R, 3, COMP, 3>), grid, block, 0, s, a); break;
            case 4: hipLaunchKernelGGL((k_stage<T, R, 4, COMP, 3>), grid, block, 0, s, a); break;
            case 5: hipLaunchKernelGGL((k_stage<T, R, 5, COMP, 3>), grid, block, 0, s, a); break;
            case 6: hipLaunchKernelGGL((k_stage<T, R, 6, COMP, 3>), grid, block, 0, s, a); break;
            case 7: hipLaunchKernelGGL((k_stage<T, R, 7, COMP, 3>), grid, block, 0, s, a); break;
            default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// periodic x ghost rows of a stage input: stored rows 0 .. kXg-1 and
// nx+kXg .. nx+2kXg-1 are copies of the stored rows of interior x rows
// nx-kXg .. nx-1 and 0 .. kXg-1 (whole stored rows, z ghost rows included),
// moved as 16-byte words
// ---------------------------------------------------------------------------
struct XgArgs {
    unsigned long long y;   // state copy
    long long words;        // 16-byte words of one copy per stored x row (nzs records)
    long long rowstride;    // 16-byte words between stored x rows
    int recw, recstride;    // 16-byte words of one site record, between site records
    int nx;
};
constexpr int kXgPerThread = 4;

XgArgs xg_args(const Geometry& g, const void* y, int esize) {
    XgArgs a;
    a.y = (unsigned long long)y;
    a.recw = (int)(g.srec * esize / 16);
    a.recstride = (int)(g.srec * esize / 16);
    a.words = (long long)g.nzs * a.recw;
    a.rowstride = g.xrow * esize / 16;
    a.nx = g.nx;
    return a;
}
long long xg_blocks(long long words) { return 2 * kXg * ((words + 256 * kXgPerThread - 1) / (256 * kXgPerThread)); }

__device__ __forceinline__ void xg_body(const XgArgs& a, long long blk) {
    const long long per = (a.words + 256 * kXgPerThread - 1) / (256 * kXgPerThread);
    const int k = (int)(blk / per);                 // ghost row 0 .. 2kXg-1
    const long long c = blk % per;
    const int dst = k < kXg ? k : a.nx + k;         // stored rows 0 .. kXg-1, nx+kXg .. nx+2kXg-1
    const int src = wrapi(dst - kXg, a.nx) + kXg;   // stored row of its interior x row
    uint4* d = reinterpret_cast<uint4*>(a.y) + dst * a.rowstride;
    const uint4* sp = reinterpret_cast<const uint4*>(a.y) + src * a.rowstride;
#pragma unroll
    for (int q = 0; q < kXgPerThread; ++q) {
        const long long n = (c * kXgPerThread + q) * 256 + threadIdx.x;
        if (n < a.words) {
            const long long o = (n / a.recw) * a.recstride + n % a.recw;   // record, word in record
            d[o] = sp[o];
        }
    }
}

__global__ __launch_bounds__(256) void k_xghosts(const XgArgs a) { xg_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_xghosts_batch(const XgArgs* __restrict__ runs, const int* __restrict__ xmap) {
    const __attribute__((address_space(4))) int* cm = (const __attribute__((address_space(4))) int*)xmap;
    const int run = cm[2 * blockIdx.x], blk = cm[2 * blockIdx.x + 1];
    const XgArgs a = runs[run];
    xg_body(a, blk);
}

template <typename T>
hipError_t launch_xghosts(const Geometry& g, T* y, hipStream_t s) {
    const XgArgs a = xg_args(g, y, sizeof(T));
    hipLaunchKernelGGL(k_xghosts, dim3((unsigned)xg_blocks(a.words)), dim3(256), 0, s, a);
    return hipGetLastError();
}

struct TotArgs {
    const double* part;
    double* tot;
    int nch;
    long long nsites;
};

template <bool COMP>
__device__ __forceinline__ void totals_body(const double* __restrict__ part, int nch, long long nsites,
                                            double* __restrict__ tot, long long n) {
    if (n >= nsites * kMomStride) return;
    const long long site = n / kMomStride;
    const int s = (int)(n % kMomStride);
    if (s >= kNMom) { tot[n] = 0.0; return; }
    const double* rec = part + site * nch * kMomRec + s;
    double sum = rec[0];
    if constexpr (!COMP) {
        for (int c = 1; c < nch; ++c) sum += rec[c * kMomRec];
        tot[n] = sum;
    } else {
        double err = 0.0;
        for (int c = 1; c < nch; ++c) {
            double s2, e;
            two_sum(sum, rec[c * kMomRec], s2, e);
            sum = s2;
            err += e;
        }
        tot[n] = sum + err;
    }
}

// ---------------------------------------------------------------------------
// per-site moment totals: sum of the nch chunk records in chunk order
// ---------------------------------------------------------------------------
template <bool COMP>
__global__ __launch_bounds__(256) void k_moment_totals(const double* __restrict__ part, int nch,
                                                       long long nsites, double* __restrict__ tot) {
    totals_body<COMP>(part, nch, nsites, tot, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

hipError_t launch_moment_totals(const Geometry& g, bool compensated, const double* part, double* tot,
                                hipStream_t s) {
    const long long nsites = (long long)g.nx * g.nz;
    const long long n = nsites * kMomStride;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (compensated) hipLaunchKernelGGL((k_moment_totals<true>), grid, block, 0, s, part, g.nch, nsites, tot);
    else hipLaunchKernelGGL((k_moment_totals<false>), grid, block, 0, s, part, g.nch, nsites, tot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// batched runs: one launch per RK4 stage (and one per moment-total pass) for
// every member run; the runs keep their own buffers, plans and physics
// ---------------------------------------------------------------------------
template <bool COMP>
__global__ __launch_bounds__(256) void k_moment_totals_batch(const TotArgs* __restrict__ runs,
                                                             const int* __restrict__ bmap) {
    const __attribute__((address_space(4))) int* cm = (const __attribute__((address_space(4))) int*)bmap;
    const int run = cm[2 * blockIdx.x], blk = cm[2 * blockIdx.x + 1];
    const TotArgs ta = runs[run];
    totals_body<COMP>(ta.part, ta.nch, ta.nsites, ta.tot, (long long)blk * 256 + threadIdx.x);
}

template <typename T>
static hipError_t batch_build_t(BatchTables& t, const std::vector<BatchMember>& ms, hipStream_t s) {
    const int n = (int)ms.size();
    std::vector<int> bmap, tmap, xmap;
    for (int r = 0; r < n; ++r)
        for (int b = 0; b < ms[r].pl.nblocks; ++b) {
            bmap.push_back(r);
            bmap.push_back(b);
        }
    for (int r = 0; r < n; ++r) {
        const long long ent = (long long)ms[r].g.nx * ms[r].g.nz * kMomStride;
        const int nb = (int)((ent + 255) / 256);
        for (int b = 0; b < nb; ++b) {
            tmap.push_back(r);
            tmap.push_back(b);
        }
    }
    for (int r = 0; r < n; ++r) {
        const long long nb = xg_blocks(xg_args(ms[r].g, nullptr, sizeof(T)).words);
        for (long long b = 0; b < nb; ++b) {
            xmap.push_back(r);
            xmap.push_back((int)b);
        }
    }
    t.nblocks = (int)bmap.size() / 2;
    t.tblocks = (int)tmap.size() / 2;
    t.xblocks = (int)xmap.size() / 2;
    hipError_t e;
    // persistent launch: one counter per XCD per stage, and a grid of as many
    // blocks as the GPU holds at once (stage 2, the largest, sets the count)
    if ((e = hipMalloc((void**)&t.ctr, kBatchCounters * sizeof(int))) != hipSuccess) return e;
    {
        int dev = 0, cus = 0, bpc = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
        constexpr int R = StageShape<T>::R;
        if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_stage_batch_q<T, R, 2, sizeof(T) == 4>,
                                                              64 * StageShape<T>::W, 0)) != hipSuccess)
            return e;
        t.pgrid = std::max(8, (cus * std::max(bpc, 1) + 7) / 8 * 8);
    }
    for (auto [dst, v] : {std::pair<int**, std::vector<int>*>{&t.bmap, &bmap}, {&t.tmap, &tmap}, {&t.xmap, &xmap}}) {
        if ((e = hipMalloc((void**)dst, v->size() * sizeof(int))) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(*dst, v->data(), v->size() * sizeof(int), hipMemcpyHostToDevice, s)) != hipSuccess)
            return e;
    }
    for (int st = 1; st <= 4; ++st) {
        std::vector<StageArgs<T>> args(n);
        std::vector<TotArgs> tots(n);
        std::vector<XgArgs> xgs(n);
        for (int r = 0; r < n; ++r) {
            const BatchMember& m = ms[r];
            const double dt = m.dt, c6 = 1.0 / 6.0 * dt, c2 = 0.5 * dt;
            T *u = (T*)m.u, *w = (T*)m.w, *ya = (T*)m.ya, *yb = (T*)m.yb;
            T* yin;
            const T* uu;
            T* yout;
            const double* mt;
            double* mo;
            double cw, cy;
            switch (st) {
                case 1: yin = u; uu = u; yout = ya; mt = m.mu; mo = m.ma; cw = 0.0; cy = c2; break;
                case 2: yin = ya; uu = u; yout = yb; mt = m.ma; mo = m.mb; cw = dt; cy = c2; break;
                case 3: yin = yb; uu = u; yout = ya; mt = m.mb; mo = m.ma; cw = 0.0; cy = dt; break;
                default: yin = ya; uu = ya; yout = u; mt = m.ma; mo = m.mu; cw = c6; cy = 0.0; break;
            }
            args[r] = make_stage_args<T>(m.g, m.ph, m.pl, yin, uu, w, yout, mt, m.mpart, cw, cy);
            tots[r] = TotArgs{m.mpart, mo, m.g.nch, (long long)m.g.nx * m.g.nz};
            xgs[r] = xg_args(m.g, yin, sizeof(T));
        }
        if ((e = hipMalloc(&t.args[st - 1], n * sizeof(StageArgs<T>))) != hipSuccess) return e;
        if ((e = hipMalloc(&t.tot[st - 1], n * sizeof(TotArgs))) != hipSuccess) return e;
        if ((e = hipMalloc(&t.xg[st - 1], n * sizeof(XgArgs))) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(t.args[st - 1], args.data(), n * sizeof(StageArgs<T>), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(t.tot[st - 1], tots.data(), n * sizeof(TotArgs), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(t.xg[st - 1], xgs.data(), n * sizeof(XgArgs), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return e;
    }
    return hipStreamSynchronize(s);
}

hipError_t batch_build(BatchTables& t, const std::vector<BatchMember>& ms, int esize, int launch, hipStream_t s) {
    batch_free(t);
    t.esize = esize;
    t.launch = launch;
    if (ms.empty()) return hipSuccess;
    for (const auto& m : ms) {
        if (!buffer_addressable(m.g, esize)) return hipErrorInvalidValue;
        if (m.g.lpr != dirs_per_chunk(esize) || m.g.wpb != waves_per_block(esize) || m.g.nox)
            return hipErrorInvalidValue;
    }
    return esize == 8 ? batch_build_t<double>(t, ms, s) : batch_build_t<float>(t, ms, s);
}

void batch_free(BatchTables& t) {
    for (int k = 0; k < 4; ++k) {
        if (t.args[k]) (void)hipFree(t.args[k]);
        if (t.tot[k]) (void)hipFree(t.tot[k]);
        if (t.xg[k]) (void)hipFree(t.xg[k]);
        t.args[k] = t.tot[k] = t.xg[k] = nullptr;
    }
    for (int** m : {&t.bmap, &t.tmap, &t.xmap, &t.ctr}) {
        if (*m) (void)hipFree(*m);
        *m = nullptr;
    }
    t.nblocks = t.tblocks = t.xblocks = t.pgrid = 0;
}

template <typename T>
static hipError_t batch_stage_t(const BatchTables& t, int st, hipStream_t s) {
    constexpr int R = StageShape<T>::R;
    constexpr bool COMP = sizeof(T) == 4;
    const dim3 grid(t.nblocks), block(64 * StageShape<T>::W);
    const StageArgs<T>* a = (const StageArgs<T>*)t.args[st - 1];
    if (t.launch == kBatchPersistent) {
        int* c = t.ctr + 8 * (st - 1);
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, c, 8);   // a kernel node, not a memset, in graphs
        const dim3 pg(t.pgrid);
        const int n = t.nblocks;
        switch (st) {
            case 1: hipLaunchKernelGGL((k_stage_batch_q<T, R, 1, COMP>), pg, block, 0, s, a, t.bmap, c, n); break;
            case 2: hipLaunchKernelGGL((k_stage_batch_q<T, R, 2, COMP>), pg, block, 0, s, a, t.bmap, c, n); break;
            case 3: hipLaunchKernelGGL((k_stage_batch_q<T, R, 3, COMP>), pg, block, 0, s, a, t.bmap, c, n); break;
            default: hipLaunchKernelGGL((k_stage_batch_q<T, R, 4, COMP>), pg, block, 0, s, a, t.bmap, c, n); break;
        }
        return hipGetLastError();
    }
    switch (st) {
        case 1: hipLaunchKernelGGL((k_stage_batch<T, R, 1, COMP>), grid, block, 0, s, a, t.bmap); break;
        case 2: hipLaunchKernelGGL((k_stage_batch<T, R, 2, COMP>), grid, block, 0, s, a, t.bmap); break;
        case 3: hipLaunchKernelGGL((k_stage_batch<T, R, 3, COMP>), grid, block, 0, s, a, t.bmap); break;
        default: hipLaunchKernelGGL((k_stage_batch<T, R, 4, COMP>), grid, block, 0, s, a, t.bmap); break;
    }
    return hipGetLastError();
}

hipError_t batch_stage(const BatchTables& t, int stage, hipStream_t s) {
    if (t.nblocks == 0) return hipSuccess;
    if (stage < 1 || stage > 4) return hipErrorInvalidValue;
    return t.esize == 8 ? batch_stage_t<double>(t, stage, s) : batch_stage_t<float>(t, stage, s);
}

// per stage: x ghost rows of every member's stage input, the stage kernel,
// the moment totals of its output
hipError_t batch_step(const BatchTables& t, hipStream_t s) {
    if (t.nblocks == 0) return hipSuccess;
    const bool comp = t.esize == 4;
    for (int st = 1; st <= 4; ++st) {
        hipLaunchKernelGGL(k_xghosts_batch, dim3(t.xblocks), dim3(256), 0, s, (const XgArgs*)t.xg[st - 1], t.xmap);
        const hipError_t e = t.esize == 8 ? batch_stage_t<double>(t, st, s) : batch_stage_t<float>(t, st, s);
        if (e != hipSuccess) return e;
        if (comp)
            hipLaunchKernelGGL((k_moment_totals_batch<true>), dim3(t.tblocks), dim3(256), 0, s,
                               (const TotArgs*)t.tot[st - 1], t.tmap);
        else
            hipLaunchKernelGGL((k_moment_totals_batch<false>), dim3(t.tblocks), dim3(256), 0, s,
                               (const TotArgs*)t.tot[st - 1], t.tmap);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// stand-alone partial moments of a state (after init / upload): one thread
// per (site, chunk), directions of the chunk summed in order
// ---------------------------------------------------------------------------
template <typename T, bool COMP>
__global__ __launch_bounds__(256) void k_moments(const Geometry g, const T* __restrict__ y,
                                                 double* __restrict__ mpart) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nrec = (long long)g.nx * g.nz * g.nch;
    if (n >= nrec) return;
    const long long site = n / g.nch;
    const int ch = (int)(n % g.nch);
    double S[kNMom], E[kNMom];
#pragma unroll
    for (int s = 0; s < kNMom; ++s) { S[s] = 0.0; E[s] = 0.0; }
    const int v1 = min(g.nv, (ch + 1) * g.lpr);
    const long long so = site_off(g, site);
    for (int v = ch * g.lpr; v < v1; ++v) {
        T yv[kNVar];
#pragma unroll
        for (int f = 0; f < kNVar; ++f) yv[f] = y[so + f * g.nvp + v];
        const double vw = g.vw[v];
        const double q[4] = {vw, g.vx[v] * vw, g.vy[v] * vw, g.vz[v] * vw};
        double t[16];
        moment_terms<T>(yv, q, t);
#pragma unroll
        for (int s = 0; s < kNMom; ++s) {
            if constexpr (COMP) {
                double s2, err;
                two_sum(S[s], t[s], s2, err);
                S[s] = s2;
                E[s] += err;
            } else {
                S[s] += t[s];
            }
        }
    }
#pragma unroll
    for (int s = 0; s < kNMom; ++s) mpart[n * kMomRec + s] = S[s] + E[s];
}

template <typename T>
hipError_t launch_moments(const Geometry& g, const T* y, double* mpart, hipStream_t s) {
    const long long nrec = (long long)g.nx * g.nz * g.nch;
    hipLaunchKernelGGL((k_moments<T, sizeof(T) == 4>), dim3((unsigned)((nrec + 255) / 256)), dim3(256), 0, s, g,
                       y, mpart);
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
    const long long o = site_off(g, site) + v;
    if (v >= g.nv) {  // row padding
        for (int f = 0; f < g.nf; ++f) y[o + f * g.nvp] = (T)0;
        return;
    }
    // 3 flavours: the tau fields keep their construction value, zero (the
    // reference's fillInitValue writes the 8 two-flavour fields only)
    for (int f = kNVar; f < g.nf; ++f) y[o + f * g.nvp] = (T)0;
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
    y[o + 0 * g.nvp] = (T)(hg * (1.0 + p3o));
    y[o + 1 * g.nvp] = (T)(hg * (1.0 - p3o));
    y[o + 2 * g.nvp] = (T)(hg * tmpr);
    y[o + 3 * g.nvp] = (T)(ipt == 5 ? hg * tmpi : 0.0);
    y[o + 4 * g.nvp] = (T)(hgb * (1.0 + p3o));
    y[o + 5 * g.nvp] = (T)(hgb * (1.0 - p3o));
    y[o + 6 * g.nvp] = (T)(hgb * tmpr);
    y[o + 7 * g.nvp] = (T)(ipt == 5 ? -hgb * tmpi : 0.0);
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
        const long long o = site_off(g, site) + v;
        double f[kNVar];
        bool finite = true;
#pragma unroll
        for (int q = 0; q < kNVar; ++q) {
            f[q] = (double)y[o + q * g.nvp];
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
__global__ void k_ghosted(const Geometry g, int gx, int gz, const T* __restrict__ y, int f,
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
    // x periodic by wrap; z periodic by wrap, or (open z) the stored ghost rows
    const int iw = wrapi(i, g.nx), jw = g.zoff ? j : wrapi(j, g.nz);
    dst[n] = (double)y[elem_off(g, iw, jw + g.zoff, f) + v];
}

template <typename T>
hipError_t launch_ghosted(const Geometry& g, int gx, int gz, const T* y, int f, double* dst, hipStream_t s) {
    const long long total = (long long)(g.nx + 2 * gx) * (g.nz + 2 * gz) * g.nv;
    hipLaunchKernelGGL((k_ghosted<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, gx, gz,
                       y, f, dst);
    return hipGetLastError();
}

template <typename T>
__global__ void k_pack(const Geometry g, const T* __restrict__ y, int f, double* __restrict__ dst) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.nz * g.nv;
    if (n >= total) return;
    const long long s = n / g.nv;
    const int v = (int)(n % g.nv);
    dst[n] = (double)y[site_off(g, s) + f * g.nvp + v];
}

template <typename T>
__global__ void k_unpack(const Geometry g, const double* __restrict__ src, T* __restrict__ y, int f) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.nz * g.nvp;
    if (n >= total) return;
    const long long s = n / g.nvp;
    const int v = (int)(n % g.nvp);
    y[site_off(g, s) + f * g.nvp + v] = (v < g.nv) ? (T)src[s * g.nv + v] : (T)0;
}

template <typename T>
hipError_t launch_pack(const Geometry& g, const T* y, int f, double* dst, hipStream_t s) {
    const long long total = (long long)g.nx * g.nz * g.nv;
    hipLaunchKernelGGL((k_pack<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, y, f, dst);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_unpack(const Geometry& g, const double* src, T* y, int f, hipStream_t s) {
    const long long total = (long long)g.nx * g.nz * g.nvp;
    hipLaunchKernelGGL((k_unpack<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, src, y, f);
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
    const long long o = elem_off(g, xi[a], j + g.zoff, 0) + v;
    const double iG = 1.0 / G0[v];  // nuosc_ana.cpp:10,14
    dst[n] = ((double)y[o + 0 * g.nvp] - (double)y[o + 1 * g.nvp]) * iG;
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
    const long long o = site_off(g, n / g.nvp) + v;
    double f[kNVar];
#pragma unroll
    for (int q = 0; q < kNVar; ++q) f[q] = (double)y[o + q * g.nvp];
    const double iG = 1.0 / G0[v], iGb = 1.0 / G0b[v];
    const double P1 = f[2] * iG, P2 = -f[3] * iG, P3 = (f[0] - f[1]) * iG;
    const double P1b = f[6] * iGb, P2b = f[7] * iGb, P3b = (f[4] - f[5]) * iGb;
    const double iP = 1.0 / sqrt(P1 * P1 + P2 * P2 + P3 * P3);
    const double iPb = 1.0 / sqrt(P1b * P1b + P2b * P2b + P3b * P3b);
    const double tmp = iP * P3 * G0[v], tmpb = iPb * P3b * G0b[v];
    y[o + 0 * g.nvp] = (T)(G0[v] + tmp);
    y[o + 1 * g.nvp] = (T)(G0[v] - tmp);
    y[o + 2 * g.nvp] = (T)(f[2] * iP);
    y[o + 3 * g.nvp] = (T)(f[3] * iP);
    y[o + 4 * g.nvp] = (T)(G0b[v] + tmpb);
    y[o + 5 * g.nvp] = (T)(G0b[v] - tmpb);
    y[o + 6 * g.nvp] = (T)(f[6] * iPb);
    y[o + 7 * g.nvp] = (T)(f[7] * iPb);
}

template <typename T>
hipError_t launch_renorm(const Geometry& g, T* y, const double* G0, const double* G0b, hipStream_t s) {
    const long long total = (long long)g.nx * g.nz * g.nvp;
    hipLaunchKernelGGL((k_renorm<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, y, G0, G0b);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// open z boundary (no reference counterpart: updateInjetOpenBoundary,
// nuosc_boundary.cpp:54-57, is a stub)
// ---------------------------------------------------------------------------
// One thread per (x row, side, ghost depth, direction) writes all fields.
template <typename T>
__global__ __launch_bounds__(256) void k_open_ghosts(const Geometry g, T* __restrict__ y,
                                                     const double* __restrict__ G0,
                                                     const double* __restrict__ G0b, const Inflow in,
                                                     const double* __restrict__ t_dev, double cdt) {
#pragma clang fp contract(off)
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * 2 * g.zoff * g.nv;
    if (n >= total) return;
    const int v = (int)(n % g.nv);
    long long r = n / g.nv;
    const int q = (int)(r % g.zoff) + 1;  // ghost depth 1..zoff
    r /= g.zoff;
    const int side = (int)(r % 2);
    const int i = (int)(r / 2);
    const int j = side == 0 ? -q : g.nz - 1 + q;
    const long long o = elem_off(g, i, j + g.zoff, 0) + v;
    const double vz = g.vz[v];
    const bool inflow = side == 0 ? (vz > 0.0) : (vz < 0.0);
    if (inflow) {
        double a = 0.0;
        if (in.on) {
            const double t = t_dev[0] + cdt;
            const double z = in.z0 + (j + 0.5) * in.dz;
            const double zz = z - vz * t - in.zc;
            a = in.eps0 * exp(-zz * zz / (2.0 * in.sigma * in.sigma));
        }
        const double p3o = sqrt(1.0 - a * a);
        const double hg = 0.5 * G0[v], hgb = 0.5 * G0b[v];
        y[o + 0 * g.nvp] = (T)(hg * (1.0 + p3o));
        y[o + 1 * g.nvp] = (T)(hg * (1.0 - p3o));
        y[o + 2 * g.nvp] = (T)(hg * a);
        y[o + 3 * g.nvp] = (T)0.0;
        y[o + 4 * g.nvp] = (T)(hgb * (1.0 + p3o));
        y[o + 5 * g.nvp] = (T)(hgb * (1.0 - p3o));
        y[o + 6 * g.nvp] = (T)(hgb * a);
        y[o + 7 * g.nvp] = (T)0.0;
        for (int f = kNVar; f < g.nf; ++f) y[o + f * g.nvp] = (T)0.0;
    } else {
        // quadratic extrapolation through the three edge cells
        const double w0 = q == 1 ? 3.0 : (q == 2 ? 6.0 : 10.0);
        const double w1 = q == 1 ? -3.0 : (q == 2 ? -8.0 : -15.0);
        const double w2 = q == 1 ? 1.0 : (q == 2 ? 3.0 : 6.0);
        const int e0 = side == 0 ? 0 : g.nz - 1, dir = side == 0 ? 1 : -1;
        const long long c0 = elem_off(g, i, e0 + g.zoff, 0) + v;
        const long long step = (long long)dir * g.srec;   // one site record along z
        for (int f = 0; f < g.nf; ++f) {
            const T* p = y + c0 + f * g.nvp;
            y[o + f * g.nvp] = (T)(w0 * (double)p[0] + w1 * (double)p[step] + w2 * (double)p[2 * step]);
        }
    }
}

template <typename T>
hipError_t launch_open_ghosts(const Geometry& g, T* y, const double* G0, const double* G0b, const Inflow& in,
                              const double* t_dev, double cdt, hipStream_t s) {
    if (g.zoff < 2 || g.zoff > 3 || g.nz < 3) return hipErrorInvalidValue;
    const long long total = (long long)g.nx * 2 * g.zoff * g.nv;
    hipLaunchKernelGGL((k_open_ghosts<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, y, G0,
                       G0b, in, t_dev, cdt);
    return hipGetLastError();
}

__global__ void k_advance_time(double* t_dev, double dt) {
    if (threadIdx.x == 0) t_dev[0] = t_dev[0] + dt;
}

hipError_t launch_advance_time(double* t_dev, double dt, hipStream_t s) {
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(64), 0, s, t_dev, dt);
    return hipGetLastError();
}

// explicit instantiations
#define NUOSC_INST(T)                                                                                   \
    template hipError_t launch_stage<T>(int, const Geometry&, const Physics&, const StagePlan&, const T*, \
                                        const T*, T*, T*, const double*, double*, double, double,        \
                                        hipStream_t, double, double);                                    \
    template hipError_t launch_moments<T>(const Geometry&, const T*, double*, hipStream_t);              \
    template hipError_t launch_init<T>(const Geometry&, int, const double*, const double*, const double*, \
                                       const double*, double, uint64_t, T*, hipStream_t);                \
    template hipError_t launch_analysis<T>(const Geometry&, const T*, const double*, const double*,      \
                                           double*, double*, hipStream_t);                               \
    template hipError_t launch_ghosted<T>(const Geometry&, int, int, const T*, int, double*, hipStream_t); \
    template hipError_t launch_pack<T>(const Geometry&, const T*, int, double*, hipStream_t);            \
    template hipError_t launch_unpack<T>(const Geometry&, const double*, T*, int, hipStream_t);          \
    template hipError_t launch_xghosts<T>(const Geometry&, T*, hipStream_t);                              \
    template hipError_t launch_p3_slice<T>(const Geometry&, const T*, const double*, const int*, int,    \
                                           const int*, int, double*, hipStream_t);                       \
    template hipError_t launch_renorm<T>(const Geometry&, T*, const double*, const double*, hipStream_t); \
    template hipError_t launch_open_ghosts<T>(const Geometry&, T*, const double*, const double*, const Inflow&, \
                                              const double*, double, hipStream_t);
NUOSC_INST(double)
NUOSC_INST(float)
#undef NUOSC_INST

}  // namespace nuosc
