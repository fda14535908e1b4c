// nbh_kernels.hip -- gfx950 kernels of the FP64 two-prime NTT product.
//
// Transform stages restate src/fpntt.jl:305-733 butterfly for butterfly (same
// reduce/mulmod schedule, same minimal-balanced twiddles), so every spectrum
// is bit-identical to the reference's.  What changes is the data movement:
// a group of consecutive stages runs on an LDS-resident tile (one HBM round
// trip per group instead of one sweep per stage), the chunk pack is fused into
// the first forward group, and the last forward group, the pointwise product
// and the first inverse group are one kernel.  The Garner unpack
// (src/fpntt.jl:806-925) becomes a per-limb parallel sum with a
// generate/propagate carry prefix (wavefront ballot + LDS across waves +
// look-back across chunks) instead of a serial 128-bit window.
//
// Every kernel takes its work from a device tile list, so the same code runs
// one product per launch or a whole heterogeneous batch per launch.

#include <hip/hip_runtime.h>

#include "nbh_field.h"
#include "nbh_internal.h"
#include "nbh_launch.h"

namespace nbh {

// ---------------------------------------------------------------------------
// DFT combines (src/fpntt.jl:310-313, 371-382, 459-492, 543-549)

__device__ __forceinline__ void dft3(const Fld &f, double &a, double &b, double &c, double w3) {
    double u = fp_mulmod3(f, b - c, w3);
    double y0 = a + (b + c);
    double y1 = (a - c) + u;
    double y2 = (a - b) - u;
    a = y0; b = y1; c = y2;
}

__device__ __forceinline__ void dft4(const Fld &f, double &a, double &b, double &c, double &d, double wi) {
    double apc = a + c, amc = a - c;
    double bpd = b + d;
    double ibmd = fp_mulmod3(f, b - d, wi);
    a = apc + bpd;
    b = amc + ibmd;
    c = apc - bpd;
    d = amc - ibmd;
}

__device__ __forceinline__ void dft5(const Fld &f, double *y, const double *k) {
    double x0 = y[0];
    double t1 = y[1] + y[4], t3 = y[1] - y[4];
    double t2 = y[2] + y[3], t4 = y[2] - y[3];
    double S = t1 + t2, D = t1 - t2;
    double y0 = x0 + S;
    double r = fp_reduce(f, x0 + fp_mulmod3(f, S, k[0]));
    double q = fp_mulmod3(f, D, k[1]);
    double g1 = fp_mulmod3(f, t3, k[2]) + fp_mulmod3(f, t4, k[3]);
    double g2 = fp_mulmod3(f, t3, k[3]) - fp_mulmod3(f, t4, k[2]);
    double u = r + q, v = r - q;
    y[0] = y0;
    y[1] = u + g1;
    y[2] = v + g2;
    y[3] = v - g2;
    y[4] = u - g1;
}

template <bool RED>
__device__ __forceinline__ void dft7(const Fld &f, double *y, const double *rot) {
    double x0 = y[0];
    double s1 = y[1] + y[6], d1 = y[1] - y[6];
    double s2 = y[2] + y[5], d2 = y[2] - y[5];
    double s3 = y[3] + y[4], d3 = y[3] - y[4];
    if (RED) {
        s1 = fp_reduce(f, s1); s2 = fp_reduce(f, s2); s3 = fp_reduce(f, s3);
        d1 = fp_reduce(f, d1); d2 = fp_reduce(f, d2); d3 = fp_reduce(f, d3);
    }
    double S = (s1 + s2) + s3;
    double y0 = x0 + S;
    double m1 = fp_mulmod3(f, S, rot[0]);
    double va = s1 - s3, vb = s2 - s3;
    double w0 = fp_mulmod3(f, va, rot[1]);
    double w1 = fp_mulmod3(f, vb, rot[2]);
    double w2 = fp_mulmod3(f, va + vb, rot[3]);
    double mn = fp_mulmod3(f, (d1 + d2) - d3, rot[4]);
    double wa = d1 + d3, wb = -(d2 + d3);
    double q0 = fp_mulmod3(f, wa, rot[5]);
    double q1 = fp_mulmod3(f, wb, rot[6]);
    double q2 = fp_mulmod3(f, wa + wb, rot[7]);
    double u1 = fp_reduce(f, m1 + (3.0 * w0 - w2));
    double u3 = fp_reduce(f, (m1 + 2.0 * w2) - 3.0 * (w0 + w1));
    double u2 = fp_reduce(f, m1 + (3.0 * w1 - w2));
    double r0 = q0 - q1, r1 = q2 - q0;
    double v1 = fp_reduce(f, mn + (2.0 * r0 + r1));
    double v3 = fp_reduce(f, (r0 + 2.0 * r1) - mn);
    double v2 = fp_reduce(f, (mn - r0) + r1);
    double a1 = x0 + u1, a2 = x0 + u2, a3 = x0 + u3;
    y[0] = y0;
    y[1] = a1 + v1;
    y[2] = a2 + v2;
    y[3] = a3 + v3;
    y[4] = a3 - v3;
    y[5] = a2 - v2;
    y[6] = a1 - v1;
}

// ---------------------------------------------------------------------------
// Uniform loads.  Plan, job and tile records live in global memory that the
// kernels also write through other pointers, so the compiler cannot prove
// their loads uniform.  readfirstlane moves them to scalar registers: every
// branch on a radix or a size becomes a scalar branch, and the index math
// derived from them runs on the scalar unit instead of the FP64 VALU.

__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniu(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)v); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
    const uint32_t lo = uniu((uint32_t)v), hi = uniu((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double unid(double v) {
    return __builtin_bit_cast(double, uni64(__builtin_bit_cast(int64_t, v)));
}
template <class T> __device__ __forceinline__ T *unip(T *p) { return (T *)uni64((int64_t)p); }

// LDS index with one pad double per 16 (see lds_doubles in nbh_internal.h):
// butterfly legs 16*k doubles apart and the lane strides of the last passes
// land in distinct banks.
__device__ __forceinline__ int pidx(int i) { return i + (i >> 4); }

// Loads and stores through the global address space: generic (flat) accesses
// would make every wait cover both the LDS and the memory counters.
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const uint64_t gcu64;
__device__ __forceinline__ double gld(const double *p, int64_t i) { return ((gcdouble *)p)[i]; }
// Spectra between passes are streamed: each pass reads its tile once and
// writes it once, in whole coalesced lines, so those accesses are marked
// nontemporal and leave the caches to the twiddle tables (measured: batch
// 3% faster).  Result limbs keep ordinary stores: their lanes write 32-byte
// pieces, and nontemporal partial lines cost more than they save.
__device__ __forceinline__ double gldn(const double *p, int64_t i) {
    return __builtin_nontemporal_load(&((gcdouble *)p)[i]);
}
__device__ __forceinline__ void gstn(double *p, int64_t i, double v) { __builtin_nontemporal_store(v, &((gdouble *)p)[i]); }
__device__ __forceinline__ uint64_t gldu(const uint64_t *p, int64_t i) { return ((gcu64 *)p)[i]; }

// Twiddle at column j of the row that starts rowoff doubles into a stage
// table.  TW32: the row start and j are both below 2^29, so the load goes
// through a raw buffer over the table with the row start as the wave-uniform
// (scalar) offset and j as the 32-bit lane offset: no 64-bit address add per
// load.  The mid tile's passes read rows and columns below 4 * (tile size),
// always in range.
template <bool TW32>
__device__ __forceinline__ double twld(const double *tw, int64_t rowoff, int64_t j) {
    if (TW32) {
        // raw buffer over the table: row start as the scalar offset, column as the lane offset
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)tw, (short)0, -1, 0x00020000);
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (uint32_t)j << 3, (uint32_t)rowoff << 3, 0);
        return __builtin_bit_cast(double, v);
    }
    return gld(tw, rowoff + j);
}

// Descriptor records (plans, jobs, tiles) are read-only while any kernel runs;
// reading them through the constant address space lets the compiler use
// scalar loads (scalar cache) instead of vector loads after every barrier.
template <class T> __device__ __forceinline__ const __attribute__((address_space(4))) T *cst(const T *p) {
    return (const __attribute__((address_space(4))) T *)p;
}

// One stage's descriptor in scalar registers.
struct StageU {
    int m, q;
    uint64_t qmagic;
    int64_t Q;
    const double *tw;
    double rot[8];
};

__device__ __forceinline__ StageU stage_u(const StageDesc *sdp) {
    const auto *sd = cst(sdp);
    StageU s;
    s.m = uni(sd->m);
    s.q = uni(sd->q);
    s.qmagic = (uint64_t)uni64((int64_t)sd->q_magic);
    s.Q = uni64(sd->Q);
    s.tw = (const double *)uni64((int64_t)sd->tw);
#pragma unroll
    for (int i = 0; i < 8; i++) s.rot[i] = unid(sd->rot[i]);
    return s;
}

// Radix-4 butterfly with its twiddles (src/fpntt.jl:645-701): forward
// combines then reduces leg 0 and twiddles legs 1..3; reverse twiddles first.
template <bool FWD, bool TW32 = false>
__device__ __forceinline__ void bfly4(const Fld &f, double &y0, double &y1, double &y2, double &y3, double wi,
                                      const double *__restrict__ tw, int64_t Q, int64_t j, int64_t rowoff = 0) {
    const double w1 = twld<TW32>(tw, rowoff, j), w2 = twld<TW32>(tw, Q + rowoff, j),
                 w3 = twld<TW32>(tw, 2 * Q + rowoff, j);
    if (FWD) dft4(f, y0, y1, y2, y3, wi);
    y0 = fp_reduce(f, y0);
    y1 = fp_mulmod3(f, y1, w1);
    y2 = fp_mulmod3(f, y2, w2);
    y3 = fp_mulmod3(f, y3, w3);
    if (!FWD) dft4(f, y0, y1, y2, y3, wi);
}

// Tile geometry shared by every pass: S rows x C columns, legs q rows apart,
// twiddle index of a butterfly = jj*Lb + jcol0 + c.  A pass may run over two
// tiles (the two operands of the mid kernel) in one loop.
struct TileGeo {
    int S, logC;
    int64_t Lb, jcol0;
};

// Two consecutive radix-4 stages s (stride q0 = 4 q1) and s+1 (stride q1) on
// 16 points per thread held in registers: the sub-problem of rows
// ob*16 q1 + jj + a q1 + b q0, a, b in 0..3.  Stage s's butterflies run over
// b (twiddle index jt + a Q1), stage s+1's over a (twiddle index jt); the
// reverse runs them in the opposite order.  Exactly the butterflies of two
// fp_radix4! sweeps, one LDS round trip instead of two.
// LIN: both leg strides are multiples of 16 points, so the padded index of
// leg (a, b) is pidx(base) + a*pidx-stride(q1) + b*pidx-stride(q0), one add
// per point instead of a pidx per point.  Otherwise the padded offsets still
// add: pidx(base + a sa + b sb) = pidx(base) + oa[a] + ob[b] with
// oa[a] = pidx(base + a sa) - pidx(base) and ob[b] likewise, because q1 is a
// power of two, sb = 4 sa and base mod 16 < sa whenever sa < 16, so the a and
// b offsets never cross a 16-point boundary together: six pidx per
// sub-problem instead of sixteen.
// CQ1 > 0: a one-column tile (logC = 0) whose pair has q1 = CQ1, so every
// offset is a constant, pidx(a q1) + pidx(4 b q1) (the decomposition above
// with base mod 16 < q1), folded into the LDS instructions.
template <bool FWD, bool LIN, bool TW32 = false, int CQ1 = 0>
__device__ void pass44(const Fld &f, double *__restrict__ x0, double *__restrict__ x1, int narr, const StageU &s0,
                       const StageU &s1, const TileGeo &g) {
    const int q1 = CQ1 ? CQ1 : s1.q;
    const int logq1 = __builtin_ctz((unsigned)q1);
    const int q0 = CQ1 ? 4 * CQ1 : s0.q;
    const int logC = CQ1 ? 0 : g.logC;
    const int C = 1 << logC;
    const int nsub = (g.S >> 4) << logC;
    for (int sg = threadIdx.x; sg < nsub * narr; sg += blockDim.x) {
        double *__restrict__ x = sg < nsub ? x0 : x1;
        const int sl = sg < nsub ? sg : sg - nsub;
        const int c = sl & (C - 1);
        const int r = sl >> logC;
        const int jj = r & (q1 - 1);
        const int row0 = ((r >> logq1) << 4) * q1 + jj;
        const int base = (row0 << logC) + c;
        const int p0 = pidx(base);
        const int sa = q1 << logC, sb = q0 << logC;
        const int pa = sa + (sa >> 4), pb = sb + (sb >> 4);
        int oa[4], ob[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            oa[k] = CQ1 ? pidx(k * CQ1) : pidx(base + k * sa) - p0;
            ob[k] = CQ1 ? pidx(4 * k * CQ1) : pidx(base + k * sb) - p0;
        }
        auto ix = [&](int a, int b) { return LIN && !CQ1 ? p0 + a * pa + b * pb : p0 + oa[a] + ob[b]; };
        double y[4][4];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) y[a][b] = x[ix(a, b)];
        const int64_t jt = (int64_t)jj * g.Lb + g.jcol0 + c;
        if (FWD) {
#pragma unroll
            for (int a = 0; a < 4; a++)
                bfly4<true, TW32>(f, y[a][0], y[a][1], y[a][2], y[a][3], s0.rot[0], s0.tw, s0.Q, jt, a * s1.Q);
#pragma unroll
            for (int b = 0; b < 4; b++) bfly4<true, TW32>(f, y[0][b], y[1][b], y[2][b], y[3][b], s1.rot[0], s1.tw, s1.Q, jt);
        } else {
#pragma unroll
            for (int b = 0; b < 4; b++)
                bfly4<false, TW32>(f, y[0][b], y[1][b], y[2][b], y[3][b], s1.rot[0], s1.tw, s1.Q, jt);
#pragma unroll
            for (int a = 0; a < 4; a++)
                bfly4<false, TW32>(f, y[a][0], y[a][1], y[a][2], y[a][3], s0.rot[0], s0.tw, s0.Q, jt, a * s1.Q);
        }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) x[ix(a, b)] = y[a][b];
    }
}

// One stage of radix M over the tile(s) (src/fpntt.jl:315-352, 384-440,
// 494-535, 630-638, 645-701): forward = combine, reduce leg 0, twiddle legs
// 1..M-1; reverse = reduce + twiddle, then combine (radix 5/7 re-reduce leg
// 0); radix 2 is twiddle-free and the same both ways.
template <bool FWD, int M, bool LIN, bool TW32 = false>
__device__ void pass1(const Fld &f, double *__restrict__ x0, double *__restrict__ x1, int narr, const StageU &s,
                      const TileGeo &g) {
    const int q = s.q;
    const int C = 1 << g.logC;
    const int nb = (g.S / M) << g.logC;
    const int strideC = q << g.logC;
    for (int bg = threadIdx.x; bg < nb * narr; bg += blockDim.x) {
        double *__restrict__ x = bg < nb ? x0 : x1;
        const int beta = bg < nb ? bg : bg - nb;
        const int c = beta & (C - 1);
        const uint32_t r = (uint32_t)beta >> g.logC;
        const uint32_t ob = (uint32_t)(((uint64_t)r * s.qmagic) >> 32);
        const uint32_t jj = r - ob * (uint32_t)q;
        const int base = ((int)(ob * (uint32_t)(M * q) + jj) << g.logC) + c;
        const int64_t j = (int64_t)jj * g.Lb + g.jcol0 + c;
        const int p0 = pidx(base), pstride = strideC + (strideC >> 4);
        auto ix = [&](int t) { return LIN ? p0 + t * pstride : pidx(base + t * strideC); };
        double y[M];
#pragma unroll
        for (int t = 0; t < M; t++) y[t] = x[ix(t)];
        if constexpr (M == 2) {
            const double u = y[0], v = y[1];
            y[0] = fp_reduce(f, u + v);
            y[1] = fp_reduce(f, u - v);
        } else if constexpr (M == 4) {
            bfly4<FWD, TW32>(f, y[0], y[1], y[2], y[3], s.rot[0], s.tw, s.Q, j);
        } else if constexpr (M == 3) {
            const double w1 = twld<TW32>(s.tw, 0, j), w2 = twld<TW32>(s.tw, s.Q, j);
            if (FWD) dft3(f, y[0], y[1], y[2], s.rot[0]);
            y[0] = fp_reduce(f, y[0]);
            y[1] = fp_mulmod3(f, y[1], w1);
            y[2] = fp_mulmod3(f, y[2], w2);
            if (!FWD) dft3(f, y[0], y[1], y[2], s.rot[0]);
        } else if constexpr (M == 5) {
            if (FWD) dft5(f, y, s.rot);
            y[0] = fp_reduce(f, y[0]);
#pragma unroll
            for (int t = 1; t < 5; t++) y[t] = fp_mulmod3(f, y[t], twld<TW32>(s.tw, (t - 1) * s.Q, j));
            if (!FWD) {
                dft5(f, y, s.rot);
                y[0] = fp_reduce(f, y[0]);
            }
        } else {
            if (FWD) dft7<false>(f, y, s.rot);
            y[0] = fp_reduce(f, y[0]);
#pragma unroll
            for (int t = 1; t < 7; t++) y[t] = fp_mulmod3(f, y[t], twld<TW32>(s.tw, (t - 1) * s.Q, j));
            if (!FWD) {
                dft7<true>(f, y, s.rot);
                y[0] = fp_reduce(f, y[0]);
            }
        }
#pragma unroll
        for (int t = 0; t < M; t++) x[ix(t)] = y[t];
    }
}

template <bool FWD, bool LIN, bool TW32>
__device__ __forceinline__ void pass_one_t(const Fld &f, double *x0, double *x1, int narr, const StageU &s, const TileGeo &g) {
    switch (s.m) {
    case 4: pass1<FWD, 4, LIN, TW32>(f, x0, x1, narr, s, g); break;
    case 2: pass1<FWD, 2, LIN, TW32>(f, x0, x1, narr, s, g); break;
    case 3: pass1<FWD, 3, LIN, TW32>(f, x0, x1, narr, s, g); break;
    case 5: pass1<FWD, 5, LIN, TW32>(f, x0, x1, narr, s, g); break;
    default: pass1<FWD, 7, LIN, TW32>(f, x0, x1, narr, s, g); break;
    }
}

template <bool FWD, bool TW32>
__device__ __forceinline__ void pass_one(const Fld &f, double *x0, double *x1, int narr, const StageU &s, const TileGeo &g) {
    if (((s.q << g.logC) & 15) == 0) pass_one_t<FWD, true, TW32>(f, x0, x1, narr, s, g);
    else pass_one_t<FWD, false, TW32>(f, x0, x1, narr, s, g);
}

template <bool FWD, bool TW32>
__device__ __forceinline__ void pass_44(const Fld &f, double *x0, double *x1, int narr, const StageU &a, const StageU &b,
                                        const TileGeo &g) {
    if (TW32) {  // the mid tile (one column): the common pair strides take constant offsets
        switch (g.logC == 0 ? b.q : 0) {
        case 2: pass44<FWD, false, TW32, 2>(f, x0, x1, narr, a, b, g); return;
        case 4: pass44<FWD, false, TW32, 4>(f, x0, x1, narr, a, b, g); return;
        case 8: pass44<FWD, false, TW32, 8>(f, x0, x1, narr, a, b, g); return;
        case 16: pass44<FWD, true, TW32, 16>(f, x0, x1, narr, a, b, g); return;
        case 32: pass44<FWD, true, TW32, 32>(f, x0, x1, narr, a, b, g); return;
        case 64: pass44<FWD, true, TW32, 64>(f, x0, x1, narr, a, b, g); return;
        default: break;
        }
    }
    if ((((a.q | b.q) << g.logC) & 15) == 0) pass44<FWD, true, TW32>(f, x0, x1, narr, a, b, g);
    else pass44<FWD, false, TW32>(f, x0, x1, narr, a, b, g);
}

// Stages [s0, s1) of a plan over the LDS tile(s), radix-4 pairs fused.  Any
// grouping of consecutive stages into passes applies the same butterflies.
// TW32: twiddle columns below 2^29 (the mid tile's; see twld).
template <bool FWD, bool TW32 = false>
__device__ void run_stages(const Fld &f, double *x0, double *x1, int narr, const DevPlan *__restrict__ P, int s0, int s1,
                           const TileGeo &g) {
    if (FWD) {
        int s = s0;
        while (s < s1) {
            const StageU a = stage_u(&P->st[s]);
            if (a.m == 4 && s + 1 < s1 && uni(cst(P)->st[s + 1].m) == 4) {
                const StageU b = stage_u(&P->st[s + 1]);
                pass_44<true, TW32>(f, x0, x1, narr, a, b, g);
                s += 2;
            } else {
                pass_one<true, TW32>(f, x0, x1, narr, a, g);
                s += 1;
            }
            __syncthreads();
        }
    } else {
        int s = s1 - 1;
        while (s >= s0) {
            const StageU b = stage_u(&P->st[s]);
            if (b.m == 4 && s - 1 >= s0 && uni(cst(P)->st[s - 1].m) == 4) {
                const StageU a = stage_u(&P->st[s - 1]);
                pass_44<false, TW32>(f, x0, x1, narr, a, b, g);
                s -= 2;
            } else {
                pass_one<false, TW32>(f, x0, x1, narr, b, g);
                s -= 1;
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Mid-kernel passes.  The forward passes below take NA arrays and load each
// twiddle once for all of them (the single-array mid kernel passes NA = 1);
// the last forward pass, the pointwise product (src/fpntt.jl:770-785) and the
// first inverse pass run on registers: one LDS round trip instead of three.
// The butterflies are exactly those of pass44 / pass1.

// x times the twiddle w^0 = 1, which is every twiddle of a transform's last
// stage (span = radix): fp_mulmod3(x, 1) = reduce(x) + fma(x, 1, -x) =
// reduce(x) + (+0.0), bit for bit, without the product and the fma.
__device__ __forceinline__ double fp_mulmod_one(const Fld &f, double x) { return fp_reduce(f, x) + 0.0; }

// One radix-M stage on registers y[0..M) with its twiddles w[0..M-2].
// UNIT: the transform's last stage, whose twiddles are all 1 (w unread).
template <bool FWD, int M, bool UNIT = false>
__device__ __forceinline__ void stage_regs(const Fld &f, double *y, const double *w, const StageU &s) {
    if constexpr (M == 2) {
        const double u = y[0], v = y[1];
        y[0] = fp_reduce(f, u + v);
        y[1] = fp_reduce(f, u - v);
    } else if constexpr (M == 4) {
        if (FWD) dft4(f, y[0], y[1], y[2], y[3], s.rot[0]);
        y[0] = fp_reduce(f, y[0]);
#pragma unroll
        for (int t = 1; t < 4; t++) y[t] = UNIT ? fp_mulmod_one(f, y[t]) : fp_mulmod3(f, y[t], w[t - 1]);
        if (!FWD) dft4(f, y[0], y[1], y[2], y[3], s.rot[0]);
    } else if constexpr (M == 3) {
        if (FWD) dft3(f, y[0], y[1], y[2], s.rot[0]);
        y[0] = fp_reduce(f, y[0]);
        y[1] = fp_mulmod3(f, y[1], w[0]);
        y[2] = fp_mulmod3(f, y[2], w[1]);
        if (!FWD) dft3(f, y[0], y[1], y[2], s.rot[0]);
    } else if constexpr (M == 5) {
        if (FWD) dft5(f, y, s.rot);
        y[0] = fp_reduce(f, y[0]);
#pragma unroll
        for (int t = 1; t < 5; t++) y[t] = fp_mulmod3(f, y[t], w[t - 1]);
        if (!FWD) {
            dft5(f, y, s.rot);
            y[0] = fp_reduce(f, y[0]);
        }
    } else {
        if (FWD) dft7<false>(f, y, s.rot);
        y[0] = fp_reduce(f, y[0]);
#pragma unroll
        for (int t = 1; t < 7; t++) y[t] = fp_mulmod3(f, y[t], w[t - 1]);
        if (!FWD) {
            dft7<true>(f, y, s.rot);
            y[0] = fp_reduce(f, y[0]);
        }
    }
}

template <int M, bool TW32 = false>
__device__ __forceinline__ void load_tw(double *w, const StageU &s, int64_t j, int64_t rowoff = 0) {
#pragma unroll
    for (int t = 1; t < M; t++) w[t - 1] = twld<TW32>(s.tw, (t - 1) * s.Q + rowoff, j);
}

// Index geometry of a pass44 sub-problem (see pass44).
struct Sub44 {
    int p0, pa, pb, base, sa, sb;
    int oa[4], ob[4];  // padded leg offsets (non-LIN form, see pass44)
    int64_t jt;
};

__device__ __forceinline__ Sub44 sub44(int sl, const StageU &s0, const StageU &s1, const TileGeo &g) {
    const int q1 = s1.q, q0 = s0.q;
    const int logq1 = __builtin_ctz((unsigned)q1);
    const int c = sl & ((1 << g.logC) - 1);
    const int r = sl >> g.logC;
    const int jj = r & (q1 - 1);
    const int row0 = ((r >> logq1) << 4) * q1 + jj;
    Sub44 u;
    u.base = (row0 << g.logC) + c;
    u.p0 = pidx(u.base);
    u.sa = q1 << g.logC;
    u.sb = q0 << g.logC;
    u.pa = u.sa + (u.sa >> 4);
    u.pb = u.sb + (u.sb >> 4);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        u.oa[k] = pidx(u.base + k * u.sa) - u.p0;
        u.ob[k] = pidx(u.base + k * u.sb) - u.p0;
    }
    u.jt = (int64_t)jj * g.Lb + g.jcol0 + c;
    return u;
}

template <bool LIN>
__device__ __forceinline__ int ix44(const Sub44 &u, int a, int b) {
    return LIN ? u.p0 + a * u.pa + b * u.pb : u.p0 + u.oa[a] + u.ob[b];
}

// Forward stages s0 (over b) and s1 (over a) of one sub-problem on registers.
template <bool FWD, bool UNIT1 = false>
__device__ __forceinline__ void regs44(const Fld &f, double (&y)[4][4], const double (&w0)[4][3], const double (&w1)[3],
                                       const StageU &s0, const StageU &s1) {
    if (FWD) {
#pragma unroll
        for (int a = 0; a < 4; a++) stage_regs<true, 4>(f, y[a], w0[a], s0);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            double t[4] = {y[0][b], y[1][b], y[2][b], y[3][b]};
            stage_regs<true, 4, UNIT1>(f, t, w1, s1);
#pragma unroll
            for (int a = 0; a < 4; a++) y[a][b] = t[a];
        }
    } else {
#pragma unroll
        for (int b = 0; b < 4; b++) {
            double t[4] = {y[0][b], y[1][b], y[2][b], y[3][b]};
            stage_regs<false, 4, UNIT1>(f, t, w1, s1);
#pragma unroll
            for (int a = 0; a < 4; a++) y[a][b] = t[a];
        }
#pragma unroll
        for (int a = 0; a < 4; a++) stage_regs<false, 4>(f, y[a], w0[a], s0);
    }
}

template <bool LIN, bool UNIT1 = false, bool TW32 = false>
__device__ __forceinline__ void tw44(double (&w0)[4][3], double (&w1)[3], const Sub44 &u, const StageU &s0,
                                     const StageU &s1) {
#pragma unroll
    for (int a = 0; a < 4; a++) load_tw<4, TW32>(w0[a], s0, u.jt, a * s1.Q);
    if (UNIT1) {
        w1[0] = w1[1] = w1[2] = 1.0;  // unread
    } else {
        load_tw<4, TW32>(w1, s1, u.jt);
    }
}

// Forward radix-4 pair on NA arrays (a, and b when NA == 2).
// CQ1 > 0: constant offsets for q1 = CQ1 on the one-column mid tile (see pass44).
template <bool LIN, int NA, int CQ1 = 0>
__device__ void pass44f(const Fld &f, double *__restrict__ xa, double *__restrict__ xb, const StageU &s0, const StageU &s1,
                        const TileGeo &g) {
    const int nsub = (g.S >> 4) << g.logC;
    StageU t1 = s1;
    TileGeo gc = g;
    if (CQ1) {
        t1.q = CQ1;
        gc.logC = 0;
    }
    for (int sl = threadIdx.x; sl < nsub; sl += blockDim.x) {
        const Sub44 u = sub44(sl, s0, t1, gc);
        double w0[4][3], w1[3];
        tw44<LIN, false, true>(w0, w1, u, s0, s1);  // the mid tile's passes only
        auto ix = [&](int a, int b) { return CQ1 ? u.p0 + pidx(a * CQ1) + pidx(4 * b * CQ1) : ix44<LIN>(u, a, b); };
#pragma unroll
        for (int n = 0; n < NA; n++) {
            double *__restrict__ x = n ? xb : xa;
            double y[4][4];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) y[a][b] = x[ix(a, b)];
            regs44<true>(f, y, w0, w1, s0, s1);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) x[ix(a, b)] = y[a][b];
        }
    }
}

// Index geometry of a single-stage butterfly (see pass1).
struct Sub1 {
    int p0, pstride, base, strideC;
    int64_t j;
};

template <int M>
__device__ __forceinline__ Sub1 sub1(int beta, const StageU &s, const TileGeo &g) {
    const int q = s.q;
    const int c = beta & ((1 << g.logC) - 1);
    const uint32_t r = (uint32_t)beta >> g.logC;
    const uint32_t ob = (uint32_t)(((uint64_t)r * s.qmagic) >> 32);
    const uint32_t jj = r - ob * (uint32_t)q;
    Sub1 u;
    u.base = ((int)(ob * (uint32_t)(M * q) + jj) << g.logC) + c;
    u.strideC = q << g.logC;
    u.p0 = pidx(u.base);
    u.pstride = u.strideC + (u.strideC >> 4);
    u.j = (int64_t)jj * g.Lb + g.jcol0 + c;
    return u;
}

template <bool LIN>
__device__ __forceinline__ int ix1(const Sub1 &u, int t) {
    return LIN ? u.p0 + t * u.pstride : pidx(u.base + t * u.strideC);
}

// Forward single stage on NA arrays.
template <int M, bool LIN, int NA>
__device__ void pass1f(const Fld &f, double *__restrict__ xa, double *__restrict__ xb, const StageU &s, const TileGeo &g) {
    const int nb = (g.S / M) << g.logC;
    for (int beta = threadIdx.x; beta < nb; beta += blockDim.x) {
        const Sub1 u = sub1<M>(beta, s, g);
        double w[M > 1 ? M - 1 : 1];
        load_tw<M, true>(w, s, u.j);  // the mid tile's passes only
#pragma unroll
        for (int n = 0; n < NA; n++) {
            double *__restrict__ x = n ? xb : xa;
            double y[M];
#pragma unroll
            for (int t = 0; t < M; t++) y[t] = x[ix1<LIN>(u, t)];
            stage_regs<true, M>(f, y, w, s);
#pragma unroll
            for (int t = 0; t < M; t++) x[ix1<LIN>(u, t)] = y[t];
        }
    }
}

template <int NA>
__device__ __forceinline__ void pass1f_any(const Fld &f, double *xa, double *xb, const StageU &s, const TileGeo &g) {
    const bool lin = ((s.q << g.logC) & 15) == 0;
    switch (s.m) {
    case 4: lin ? pass1f<4, true, NA>(f, xa, xb, s, g) : pass1f<4, false, NA>(f, xa, xb, s, g); break;
    case 2: lin ? pass1f<2, true, NA>(f, xa, xb, s, g) : pass1f<2, false, NA>(f, xa, xb, s, g); break;
    case 3: lin ? pass1f<3, true, NA>(f, xa, xb, s, g) : pass1f<3, false, NA>(f, xa, xb, s, g); break;
    case 5: lin ? pass1f<5, true, NA>(f, xa, xb, s, g) : pass1f<5, false, NA>(f, xa, xb, s, g); break;
    default: lin ? pass1f<7, true, NA>(f, xa, xb, s, g) : pass1f<7, false, NA>(f, xa, xb, s, g); break;
    }
}

// The mid kernel keeps ONE array in LDS, so twice as many product tiles share
// a CU as with both operands resident.  A product transforms b first and
// keeps the result of b's last forward pass in registers (the thread's own
// index set of that pass: one 16-point sub-problem, or 16/M butterflies);
// then a is loaded into the same array, transformed, and its last forward
// pass, the pointwise product against the held values and the same pass
// inverted run on registers.  A square multiplies a by itself there; a
// product by a cached spectrum reads b's points from HBM there.
enum MidSrc { SRC_SELF = 0, SRC_HELD = 1, SRC_GLOB = 2 };
// Held values per thread: 16 for a radix-4 pair, M * (16/M) for one radix-M
// stage.  The last stage of every admissible length m*2^k is radix 4 or 2
// (k >= 2 unless m = 1, src/fpntt.jl:236-271), so the last pass is one of
// those.
constexpr int HOLD = 16;

// Forward passes from s0 up to (not including) the last pass of [s0, s1), on
// one array; returns the first stage of the last pass.
__device__ int fwd_until_last(const Fld &f, double *x, const DevPlan *__restrict__ P, int s0, int s1, const TileGeo &g) {
    int s = s0;
    for (;;) {
        const StageU a = stage_u(&P->st[s]);
        const int m1 = s + 1 < s1 ? uni(cst(P)->st[s + 1].m) : 0;
        const bool pair = a.m == 4 && m1 == 4;
        const int e = s + (pair ? 2 : 1);
        if (e >= s1) return s;
        // a radix-4 stage followed by the final radix 2 is the last pass too
        if (a.m == 4 && m1 == 2 && s + 2 == s1) return s;
        if (pair) {
            const StageU b = stage_u(&P->st[s + 1]);
            // the mid tile is one column: the common pair strides take constant offsets
            switch (g.logC == 0 ? b.q : 0) {
            case 2: pass44f<false, 1, 2>(f, x, nullptr, a, b, g); break;
            case 4: pass44f<false, 1, 4>(f, x, nullptr, a, b, g); break;
            case 8: pass44f<false, 1, 8>(f, x, nullptr, a, b, g); break;
            case 16: pass44f<true, 1, 16>(f, x, nullptr, a, b, g); break;
            case 32: pass44f<true, 1, 32>(f, x, nullptr, a, b, g); break;
            case 64: pass44f<true, 1, 64>(f, x, nullptr, a, b, g); break;
            default:
                if ((((a.q | b.q) << g.logC) & 15) == 0) pass44f<true, 1>(f, x, nullptr, a, b, g);
                else pass44f<false, 1>(f, x, nullptr, a, b, g);
            }
        } else {
            pass1f_any<1>(f, x, nullptr, a, g);
        }
        __syncthreads();
        s = e;
    }
}

// The mid tile's last pass holds the transform's last stages on a
// contiguous tile (one column): the last stage's legs are adjacent (q = 1)
// and the stage before it has q = 4 (radix 4) or 2 (radix 2 last), so a
// sub-problem's points are one run starting at a multiple of its size, which
// never crosses a pad: constant offsets from pidx(base), which the compiler
// folds into the LDS instructions' offset fields.  LIN = false is that form.
template <bool LIN>
__device__ __forceinline__ int ix_last44(const Sub44 &u, int a, int b) {
    return LIN ? ix44<true>(u, a, b) : u.p0 + a + 4 * b;
}

// Last pass = a radix-4 pair.  HOLDB: compute x's last forward pass into hb
// (no store).  Otherwise: the fused pass on x with b from SRC.
template <bool LIN, bool HOLDB, int SRC>
__device__ void last44(const Fld &f, double *__restrict__ x, double (&hb)[HOLD], const double *__restrict__ gb, int64_t gbase,
                       const StageU &s0, const StageU &s1, const TileGeo &g) {
    const int nsub = (g.S >> 4) << g.logC;
    const int sl = threadIdx.x;  // one sub-problem per thread: blockDim >= S/16
    if (sl >= nsub) return;
    const Sub44 u = sub44(sl, s0, s1, g);
    double y[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) y[a][b] = x[ix_last44<LIN>(u, a, b)];
    // s1 is the transform's last stage: its twiddles are all 1
    double w0[4][3], w1[3];
    tw44<LIN, true, true>(w0, w1, u, s0, s1);
    regs44<true, true>(f, y, w0, w1, s0, s1);
    if (HOLDB) {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) hb[4 * a + b] = y[a][b];
        return;
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const double v = SRC == SRC_SELF ? y[a][b]
                             : SRC == SRC_HELD ? hb[4 * a + b]
                                               : gld(gb, gbase + u.base + a * u.sa + b * u.sb);
            y[a][b] = fp_mulmod3(f, y[a][b], v);
        }
    regs44<false, true>(f, y, w0, w1, s0, s1);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) x[ix_last44<LIN>(u, a, b)] = y[a][b];
}

// Last pass = one radix-M stage: the thread's butterflies beta = t + it*T.
template <int M, bool LIN, bool HOLDB, int SRC>
__device__ void last1(const Fld &f, double *__restrict__ x, double (&hb)[HOLD], const double *__restrict__ gb, int64_t gbase,
                      const StageU &s, const TileGeo &g) {
    static_assert(16 % M == 0, "the last pass is radix 4 or 2");
    constexpr int MAXIT = 16 / M;  // blockDim >= S/16, so nb/blockDim <= 16/M
    const int nb = (g.S / M) << g.logC;
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        const int beta = threadIdx.x + it * (int)blockDim.x;
        if (beta >= nb) break;
        const Sub1 u = sub1<M>(beta, s, g);
        // s is the transform's last stage: its twiddles are all 1 (w unread)
        const double w[M > 1 ? M - 1 : 1] = {};
        double y[M];
        // not LIN: the mid tile's last stage, q = 1, M-point runs at
        // multiples of M (see ix_last44)
        auto ix = [&](int t) { return LIN ? ix1<true>(u, t) : u.p0 + t; };
#pragma unroll
        for (int t = 0; t < M; t++) y[t] = x[ix(t)];
        stage_regs<true, M, true>(f, y, w, s);
        if (HOLDB) {
#pragma unroll
            for (int t = 0; t < M; t++) hb[it * M + t] = y[t];
            continue;
        }
#pragma unroll
        for (int t = 0; t < M; t++) {
            const double v = SRC == SRC_SELF ? y[t]
                             : SRC == SRC_HELD ? hb[it * M + t]
                                               : gld(gb, gbase + u.base + t * u.strideC);
            y[t] = fp_mulmod3(f, y[t], v);
        }
        stage_regs<false, M, true>(f, y, w, s);
#pragma unroll
        for (int t = 0; t < M; t++) x[ix(t)] = y[t];
    }
}

// Last pass = a radix-4 stage s0 then the transform's final radix-2 stage
// s1 (lengths with an odd power of two): 8-point sub-problems, rows
// row0 + a q1 + b q0 (a in 0..1, b in 0..3), two per thread.  s0's
// butterflies run over b (twiddle index jt + a Q1), s1's over a (no
// twiddles); the inverse runs them in the opposite order.  Exactly the
// butterflies of pass1<4> and pass1<2>, with one LDS round trip instead of
// two.
template <bool LIN, bool HOLDB, int SRC>
__device__ void last42(const Fld &f, double *__restrict__ x, double (&hb)[HOLD], const double *__restrict__ gb, int64_t gbase,
                       const StageU &s0, const StageU &s1, const TileGeo &g) {
    const int nsub = (g.S >> 3) << g.logC;
    const int q1 = s1.q, q0 = s0.q;
    const int logq1 = __builtin_ctz((unsigned)q1);
    const int cmask = (1 << g.logC) - 1;
    const int sa = q1 << g.logC, sb = q0 << g.logC;
    const int pa = sa + (sa >> 4), pb = sb + (sb >> 4);
    const double wnone[1] = {};
#pragma unroll
    for (int it = 0; it < 2; it++) {  // blockDim >= S/16: at most two sub-problems per thread
        const int sl = (int)threadIdx.x + it * (int)blockDim.x;
        if (sl >= nsub) break;
        const int c = sl & cmask;
        const int r = sl >> g.logC;
        const int jj = r & (q1 - 1);
        const int row0 = ((r >> logq1) << 3) * q1 + jj;
        const int base = (row0 << g.logC) + c;
        const int p0 = pidx(base);
        // not LIN: the mid tile's last pass, legs q1 = 1 and q0 = 2 apart in
        // 8-point runs at multiples of 8 (see ix_last44)
        auto ix = [&](int a, int b) { return LIN ? p0 + a * pa + b * pb : p0 + a + 2 * b; };
        const int64_t jt = (int64_t)jj * g.Lb + g.jcol0 + c;
        double y[2][4], w0[2][3];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) y[a][b] = x[ix(a, b)];
#pragma unroll
        for (int a = 0; a < 2; a++) load_tw<4, true>(w0[a], s0, jt, a * s1.Q);
#pragma unroll
        for (int a = 0; a < 2; a++) stage_regs<true, 4>(f, y[a], w0[a], s0);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            double t[2] = {y[0][b], y[1][b]};
            stage_regs<true, 2>(f, t, wnone, s1);
            y[0][b] = t[0];
            y[1][b] = t[1];
        }
        if (HOLDB) {
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) hb[8 * it + 4 * a + b] = y[a][b];
            continue;
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const double v = SRC == SRC_SELF ? y[a][b]
                                 : SRC == SRC_HELD ? hb[8 * it + 4 * a + b]
                                                   : gld(gb, gbase + base + a * sa + b * sb);
                y[a][b] = fp_mulmod3(f, y[a][b], v);
            }
#pragma unroll
        for (int b = 0; b < 4; b++) {
            double t[2] = {y[0][b], y[1][b]};
            stage_regs<false, 2>(f, t, wnone, s1);
            y[0][b] = t[0];
            y[1][b] = t[1];
        }
#pragma unroll
        for (int a = 0; a < 2; a++) stage_regs<false, 4>(f, y[a], w0[a], s0);
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) x[ix(a, b)] = y[a][b];
    }
}

// The mid tile's last pass: its last one or two stages, which are the
// transform's last, so their strides are below 16: the LIN = false forms,
// which index unit-stride runs (ix_last44).
template <bool HOLDB, int SRC>
__device__ __forceinline__ void last_pass(const Fld &f, double *x, double (&hb)[HOLD], const double *gb, int64_t gbase,
                                          const DevPlan *__restrict__ P, int s, int s1, const TileGeo &g) {
    const StageU a = stage_u(&P->st[s]);
    if (s1 - s == 2) {
        const StageU b = stage_u(&P->st[s + 1]);
        if (b.m == 2) last42<false, HOLDB, SRC>(f, x, hb, gb, gbase, a, b, g);
        else last44<false, HOLDB, SRC>(f, x, hb, gb, gbase, a, b, g);
        return;
    }
    if (a.m == 2) last1<2, false, HOLDB, SRC>(f, x, hb, gb, gbase, a, g);
    else last1<4, false, HOLDB, SRC>(f, x, hb, gb, gbase, a, g);
}

// Points each thread moves per round in the tile load loops.
constexpr int LU = 16;

// Chunk i of an operand: bits [i*b, i*b + b) (src/fpntt.jl:744-768); chunks
// past nch are the zero fill.  Chunks are < 2^b < p: exact, canonical.
// Branch-free: the caller passes bit = i*b and in = (i < nch); a lane that is
// not `in` reads limb 0 and yields 0, and the second limb read is clamped to
// the operand.  The chunk (< 2^49) converts exactly as hi*2^32 + lo.
__device__ __forceinline__ double pack_bits(const uint64_t *__restrict__ limbs, int64_t n, uint64_t bmask, bool in,
                                            int64_t bit) {
    bit = in ? bit : 0;
    const int64_t w = bit >> 6;
    const uint32_t sh = (uint32_t)bit & 63u;
    const bool two = w + 1 < n;
    const uint64_t lo = gldu(limbs, w);
    const uint64_t hi = gldu(limbs, two ? w + 1 : w);
    // (hi << 1) << (63 - sh) is hi << (64 - sh) for sh > 0 and 0 for sh == 0
    const uint64_t c = ((lo >> sh) | (two ? (hi << 1) << (63u - sh) : 0ull)) & bmask;
    const double d = __builtin_fma((double)(uint32_t)(c >> 32), 4294967296.0, (double)(uint32_t)c);
    return in ? d : 0.0;
}

__device__ __forceinline__ double pack_chunk(const uint64_t *__restrict__ limbs, int64_t n, int b,
                                             int64_t nch, int64_t i) {
    return pack_bits(limbs, n, (1ull << b) - 1, i < nch, i * (int64_t)b);
}

// Operands of at most this many limbs (2 GiB) are read through a buffer
// resource: 32-bit byte offsets, and a dword past the operand reads as zero
// (range checking is per dword), as the bits past the operand are.
constexpr int64_t BUF_MAX_LIMBS = (int64_t)1 << 28;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t limb_rsrc(const uint64_t *limbs, int64_t n) {
    // word 3 of a gfx9-family raw buffer descriptor: 32-bit data, no swizzle
    return __builtin_amdgcn_make_buffer_rsrc((void *)limbs, (short)0, (int)(8 * n), 0x00020000);
}

// The chunk at `bit` through the buffer: a chunk of at most 49 bits starting
// at bit (bit mod 32) of its first dword lies in that dword and the next two,
// and two funnel shifts extract it.  A lane that is not `in` reads far past
// the operand and yields 0.
__device__ __forceinline__ double pack_buf(__amdgpu_buffer_rsrc_t r, uint32_t lmask, uint32_t hmask, bool in,
                                           int64_t bit) {
    const uint32_t off = in ? (uint32_t)(bit >> 5) << 2 : 0x80000000u;
    const uint32_t sh = (uint32_t)bit & 31u;
    const uint32_t x0 = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    const uint32_t x1 = __builtin_amdgcn_raw_buffer_load_b32(r, off + 4, 0, 0);
    const uint32_t x2 = __builtin_amdgcn_raw_buffer_load_b32(r, off + 8, 0, 0);
    const uint32_t lo = __builtin_amdgcn_alignbit(x1, x0, sh) & lmask;
    const uint32_t hi = __builtin_amdgcn_alignbit(x2, x1, sh) & hmask;
    return __builtin_fma((double)hi, 4294967296.0, (double)lo);
}

// Where a tile's points come from: spectra x, or (pack) the b-bit chunks of
// an operand's limbs.
struct PointSrc {
    const double *x;
    const uint64_t *limbs;
    int64_t n, nch;
    int b;
};

// One round of NU loads per thread, all issued before any result is used:
// tile point e = e0 + k*T is transform element gbase + (e >> logC)*Lb +
// (e & (C-1)).  In a full round (every e < total) with T a multiple of C the
// element and bit positions advance by a uniform step; otherwise a lane past
// `total` re-reads point 0.
// Spectra points of a tile through a raw buffer based at the tile's first
// point: the per-lane offset is the lane's first point, the per-load offset a
// wave-uniform multiple of the round step, both 32-bit byte offsets (a mid
// tile is at most 16384 points), so no 64-bit address add per load.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const double *x) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, -1, 0x00020000);
}
__device__ __forceinline__ double bufldn(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));  // nt
}
__device__ __forceinline__ void bufstn(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double v) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, voff, soff, 2);  // nt
}

// OFF32: the mid tile (Lb = 1, one column), whose spectra loads take the
// buffer form above.
template <bool PACK, int NU, bool OFF32 = false>
__device__ __forceinline__ void load_round(double (&v)[NU], const PointSrc &src, int e0, int total, bool full,
                                           const TileGeo &g, int64_t gbase) {
    const int T = blockDim.x;
    const int cmask = (1 << g.logC) - 1;
    if (!PACK && OFF32 && full) {
        const __amdgpu_buffer_rsrc_t r = tile_rsrc(src.x + gbase);
#pragma unroll
        for (int k = 0; k < NU; k++) v[k] = bufldn(r, (uint32_t)e0 << 3, (uint32_t)(k * T) << 3);
        return;
    }
    if (full && (T & cmask) == 0) {
        const int64_t i0 = gbase + (int64_t)(e0 >> g.logC) * g.Lb + (e0 & cmask);
        const int64_t step = (int64_t)(T >> g.logC) * g.Lb;
        if (PACK) {
            const int64_t bit0 = i0 * src.b, bstep = step * src.b;
            if (src.n <= BUF_MAX_LIMBS) {
                const __amdgpu_buffer_rsrc_t r = limb_rsrc(src.limbs, src.n);
                const uint32_t lmask = src.b >= 32 ? ~0u : (1u << src.b) - 1u;
                const uint32_t hmask = src.b >= 32 ? (1u << (src.b - 32)) - 1u : 0u;
#pragma unroll
                for (int k = 0; k < NU; k++)
                    v[k] = pack_buf(r, lmask, hmask, i0 + k * step < src.nch, bit0 + k * bstep);
            } else {
                const uint64_t bmask = (1ull << src.b) - 1;
#pragma unroll
                for (int k = 0; k < NU; k++)
                    v[k] = pack_bits(src.limbs, src.n, bmask, i0 + k * step < src.nch, bit0 + k * bstep);
            }
        } else {
#pragma unroll
            for (int k = 0; k < NU; k++) v[k] = gldn(src.x, i0 + k * step);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NU; k++) {
            const int e = e0 + k * T;
            const int ec = e < total ? e : 0;
            const int64_t i = gbase + (int64_t)(ec >> g.logC) * g.Lb + (ec & cmask);
            v[k] = PACK ? pack_chunk(src.limbs, src.n, src.b, src.nch, i) : gld(src.x, i);
        }
    }
}

// Stores of a round into the padded LDS tile: T is a multiple of 64, so
// pidx(e0 + k*T) = pidx(e0) + k*(T + T/16).
template <int NU>
__device__ __forceinline__ void lds_round(double *__restrict__ lds, const double (&v)[NU], int e0, int total,
                                          bool full) {
    const int T = blockDim.x;
    const int p0 = pidx(e0), pstep = T + (T >> 4);
    if (full) {
#pragma unroll
        for (int k = 0; k < NU; k++) lds[p0 + k * pstep] = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < NU; k++)
            if (e0 + k * T < total) lds[p0 + k * pstep] = v[k];
    }
}

// A whole tile (total points) from src into LDS: full rounds of NU points per
// thread, then full rounds of 4, then single points (the last one guarded),
// so small tiles issue no redundant loads.
template <int NU, bool PACK>
__device__ __forceinline__ void load_tile_t(double *__restrict__ lds, const PointSrc &src, int total,
                                            const TileGeo &g, int64_t gbase) {
    const int T = blockDim.x;
    int r0 = 0;
    for (; r0 + NU * T <= total; r0 += NU * T) {
        double v[NU];
        load_round<PACK, NU>(v, src, r0 + (int)threadIdx.x, total, true, g, gbase);
        lds_round<NU>(lds, v, r0 + (int)threadIdx.x, total, true);
    }
    for (; r0 + 4 * T <= total; r0 += 4 * T) {
        double v[4];
        load_round<PACK, 4>(v, src, r0 + (int)threadIdx.x, total, true, g, gbase);
        lds_round<4>(lds, v, r0 + (int)threadIdx.x, total, true);
    }
    for (; r0 < total; r0 += T) {
        double v[1];
        const bool full = r0 + T <= total;
        load_round<PACK, 1>(v, src, r0 + (int)threadIdx.x, total, full, g, gbase);
        lds_round<1>(lds, v, r0 + (int)threadIdx.x, total, full);
    }
}

template <int NU>
__device__ __forceinline__ void load_tile(double *__restrict__ lds, const PointSrc &src, bool pack, int total,
                                          const TileGeo &g, int64_t gbase) {
    if (pack) load_tile_t<NU, true>(lds, src, total, g, gbase);
    else load_tile_t<NU, false>(lds, src, total, g, gbase);
}

// The tile a block works on: from the launch's tile list, or generated.
// The job of a tile: from the launch's job array, or the record the launch
// carries in its TileGen when it covers one job (no job array).
__device__ __forceinline__ const DevJob *job_of(const DevJob *jobs, const Tile &t, const TileGen &gen) {
    return jobs ? jobs + t.job : &gen.job;
}

__device__ __forceinline__ Tile fetch_tile(const Tile *__restrict__ tiles, const TileGen &g) {
    Tile t;
    if (tiles) {
        // Tile is four words: job, (prime, op, group, flags), tile, pad
        const auto *v = cst((const uint32_t *)(tiles + blockIdx.x));
        t.job = uniu(v[0]);
        const uint32_t w = uniu(v[1]);
        t.prime = (uint8_t)w;
        t.op = (uint8_t)(w >> 8);
        t.group = (uint8_t)(w >> 16);
        t.flags = (uint8_t)(w >> 24);
        t.tile = uniu(v[2]);
        t.pad = 0;
        return t;
    }
    int64_t b = blockIdx.x;
    t.job = 0;
    t.tile = (uint32_t)(b % g.per);
    b /= g.per;
    t.op = (uint8_t)(b % g.nops);
    t.prime = (uint8_t)(g.prime0 + b / g.nops);
    t.group = (uint8_t)g.group;
    t.flags = (uint8_t)g.flags;
    t.pad = 0;
    return t;
}

// The LDS tile (rows of 2^logC columns, row r at gbase + r*Lb) back to HBM,
// SU points per thread per round.  When the block's thread count is a
// multiple of the row width (always, for the schedules built here), point
// e0 + k*T of thread e0 sits k*step elements past the thread's first point
// and k*(T + T/16) doubles past its first padded LDS slot: one index per
// thread, then pointer steps by uniform amounts.
constexpr int SU = 8;

__device__ __forceinline__ void store_tile(double *__restrict__ x, const double *__restrict__ lds, int total,
                                           const TileGeo &g, int64_t gbase) {
    const int T = blockDim.x;
    const int cmask = (1 << g.logC) - 1;
    const int e0 = threadIdx.x;
    if ((T & cmask) == 0) {
        gdouble *__restrict__ p = (gdouble *)x + (gbase + (int64_t)(e0 >> g.logC) * g.Lb + (e0 & cmask));
        const int64_t step = (int64_t)(T >> g.logC) * g.Lb;
        const double *__restrict__ l = lds + pidx(e0);
        const int lstep = T + (T >> 4);
        int r0 = 0;
        for (; r0 + SU * T <= total; r0 += SU * T) {
            double v[SU];
#pragma unroll
            for (int k = 0; k < SU; k++) v[k] = l[k * lstep];
#pragma unroll
            for (int k = 0; k < SU; k++) __builtin_nontemporal_store(v[k], &p[k * step]);
            p += SU * step;
            l += SU * lstep;
        }
        for (int k = 0; r0 + k * T < total; k++)
            if (r0 + k * T + e0 < total) __builtin_nontemporal_store(l[k * lstep], &p[k * step]);
        return;
    }
    for (int r0 = 0; r0 < total; r0 += T) {
        const int e = r0 + e0;
        if (e < total) gstn(x, gbase + (int64_t)(e >> g.logC) * g.Lb + (e & cmask), lds[pidx(e)]);
    }
}

// ---------------------------------------------------------------------------
// Edge passes.  When a tile's first two stages are radix 4 and the block has
// exactly one thread per 16 points (T = S*C/16), thread e0's round loads
// e0 + k*T (k = 0..15) are rows r + k*S/16 of column c (r = e0 >> logC,
// c = e0 & (C-1)): exactly the sub-problem (a, b), k = a + 4b, of the pair's
// pass44 with sl = e0 (legs S/16 and S/4 rows apart).  So the first forward
// pair runs on the loaded registers and writes LDS once, and the inverse of
// that pair -- the last inverse pass -- stores from registers to the same
// addresses the loads read.  One LDS round trip and one barrier fewer on each
// side; the butterflies are the pass44 ones.

__device__ __forceinline__ bool edge_ok(const DevPlan *__restrict__ P, int s0, int s1, int total, int min_stages) {
    return total == 16 * (int)blockDim.x && s1 - s0 >= min_stages && uni(cst(P)->st[s0].m) == 4 &&
           uni(cst(P)->st[s0 + 1].m) == 4;
}

// Loads the thread's 16 points from src (rounds of NU), runs stages s0, s0+1
// forward on them and writes them to the LDS tile.  TW32: twiddles through
// 32-bit offsets (see twld); MIDIO: the one-column mid tile's spectra through
// a tile-based buffer (see load_round).
template <int NU, bool TW32 = false, bool MIDIO = false>
__device__ __forceinline__ void edge_load44(const Fld &f, double *__restrict__ lds, const PointSrc &src, bool pack,
                                            const TileGeo &g, int64_t gbase, const DevPlan *__restrict__ P, int s0) {
    static_assert(16 % NU == 0, "rounds of NU points cover the 16");
    const int T = blockDim.x;
    const int total = 16 * T;
    double y[4][4];
#pragma unroll
    for (int r = 0; r < 16 / NU; r++) {
        double v[NU];
        if (pack) load_round<true, NU>(v, src, (int)threadIdx.x + r * NU * T, total, true, g, gbase);
        else load_round<false, NU, MIDIO>(v, src, (int)threadIdx.x + r * NU * T, total, true, g, gbase);
#pragma unroll
        for (int k = 0; k < NU; k++) y[(r * NU + k) & 3][(r * NU + k) >> 2] = v[k];
    }
    const StageU sa = stage_u(&P->st[s0]), sb = stage_u(&P->st[s0 + 1]);
    const Sub44 u = sub44((int)threadIdx.x, sa, sb, g);
    double w0[4][3], w1[3];
    tw44<false, false, TW32>(w0, w1, u, sa, sb);
    regs44<true>(f, y, w0, w1, sa, sb);
    if ((((sa.q | sb.q) << g.logC) & 15) == 0) {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) lds[ix44<true>(u, a, b)] = y[a][b];
    } else {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) lds[ix44<false>(u, a, b)] = y[a][b];
    }
}

// Reads the thread's pass44 sub-problem of stages s0, s0+1 from the LDS tile,
// runs the pair inverse and stores the 16 points to x.
template <bool TW32 = false, bool MIDIO = false>
__device__ __forceinline__ void edge_store44(const Fld &f, double *__restrict__ x, const double *__restrict__ lds,
                                             const TileGeo &g, int64_t gbase, const DevPlan *__restrict__ P, int s0) {
    const int T = blockDim.x;
    const int e0 = threadIdx.x;
    const StageU sa = stage_u(&P->st[s0]), sb = stage_u(&P->st[s0 + 1]);
    const Sub44 u = sub44(e0, sa, sb, g);
    double y[4][4];
    if ((((sa.q | sb.q) << g.logC) & 15) == 0) {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) y[a][b] = lds[ix44<true>(u, a, b)];
    } else {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) y[a][b] = lds[ix44<false>(u, a, b)];
    }
    double w0[4][3], w1[3];
    tw44<false, false, TW32>(w0, w1, u, sa, sb);
    regs44<false>(f, y, w0, w1, sa, sb);
    if (MIDIO) {  // the mid tile: points e0 + k*T of the tile (see load_round)
        const __amdgpu_buffer_rsrc_t r = tile_rsrc(x + gbase);
#pragma unroll
        for (int k = 0; k < 16; k++) bufstn(r, (uint32_t)e0 << 3, (uint32_t)(k * T) << 3, y[k & 3][k >> 2]);
        return;
    }
    const int cmask = (1 << g.logC) - 1;
    gdouble *__restrict__ p = (gdouble *)x + (gbase + (int64_t)(e0 >> g.logC) * g.Lb + (e0 & cmask));
    const int64_t step = (int64_t)(T >> g.logC) * g.Lb;
#pragma unroll
    for (int k = 0; k < 16; k++) __builtin_nontemporal_store(y[k & 3][k >> 2], &p[k * step]);
}

// ---------------------------------------------------------------------------
// Group pass: one tile of one group of one transform, forward or inverse.

template <int NU>
__device__ __forceinline__ void group_body(const Fld &f, const DevJob *__restrict__ Jp, const Tile &t, bool fwd, double *lds) {
    const auto *J = cst(Jp);
    const DevPlan *__restrict__ P = unip(J->plan[t.prime]);
    const auto *G = &cst(P)->g[t.group];
    TileGeo g;
    g.S = uni(G->S);
    g.logC = uni(G->logC);
    g.Lb = uni64(G->Lb);
    const int64_t La = uni64(G->La);
    const int s0 = uni(G->s0), s1 = uni(G->s1);
    double *__restrict__ x = unip(t.op ? J->xb[t.prime] : J->xa[t.prime]);
    const int64_t tpr = g.Lb >> g.logC;
    const int64_t blk = (int64_t)t.tile / tpr;
    const int64_t jc = (int64_t)t.tile - blk * tpr;
    g.jcol0 = jc << g.logC;
    const int64_t gbase = blk * La + g.jcol0;
    const int total = g.S << g.logC;

    // Every thread issues all of its loads before its first LDS store so the
    // load latencies overlap (LU points per thread per round).
    PointSrc src;
    src.x = x;
    src.limbs = unip(t.op ? J->b : J->a);
    src.n = uni64(t.op ? J->n : J->m);
    src.nch = uni64(t.op ? J->ncb : J->nca);
    src.b = uni(J->bch);
    const bool pack = (t.flags & TILE_PACK) != 0;
    const bool edge = edge_ok(P, s0, s1, total, 2);
    if (fwd) {
        if (edge) {
            edge_load44<(NU < 8 ? NU : 8)>(f, lds, src, pack, g, gbase, P, s0);
            __syncthreads();
            run_stages<true>(f, lds, nullptr, 1, P, s0 + 2, s1, g);
        } else {
            load_tile<NU>(lds, src, pack, total, g, gbase);
            __syncthreads();
            run_stages<true>(f, lds, nullptr, 1, P, s0, s1, g);
        }
        store_tile(x, lds, total, g, gbase);
    } else {
        load_tile<NU>(lds, src, pack, total, g, gbase);
        __syncthreads();
        if (edge) {
            run_stages<false>(f, lds, nullptr, 1, P, s0 + 2, s1, g);
            edge_store44(f, x, lds, g, gbase, P, s0);
        } else {
            run_stages<false>(f, lds, nullptr, 1, P, s0, s1, g);
            store_tile(x, lds, total, g, gbase);
        }
    }
}

// Forward and inverse group passes are separate kernels (one code path each,
// and separate names in a kernel trace).
// NU: points each thread loads per round (see load_tile).
template <bool FWD, int NU>
__device__ __forceinline__ void group_kernel(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles,
                                             const TileGen &gen, double *lds) {
    const Tile t = fetch_tile(tiles, gen);
    const DevJob *J = job_of(jobs, t, gen);
    if (uni(cst(J)->zero)) return;
    group_body<NU>(fld(t.prime), J, t, FWD, lds);
}

__global__ void __launch_bounds__(1024) k_group_fwd(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles,
                                                    TileGen gen) {
    extern __shared__ double lds[];
    group_kernel<true, LU>(jobs, tiles, gen, lds);
}

__global__ void __launch_bounds__(1024) k_group_inv(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles,
                                                    TileGen gen) {
    extern __shared__ double lds[];
    group_kernel<false, LU>(jobs, tiles, gen, lds);
}

// The same kernels held to six waves per SIMD (at most 80 VGPRs, no spills)
// for launches whose tiles are small enough in LDS that more than the four
// waves per SIMD of the unrestricted build fit on a CU: those are bound by
// memory latency, and two more waves per SIMD hide more of it.  Launches of
// large tiles are LDS-bound at four waves and keep the unrestricted build.
// With more waves in flight, rounds of 8 loads per thread (instead of 16)
// measured faster here, and slower in the large-tile build.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6)))
k_group_fwd_w6(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles, TileGen gen) {
    extern __shared__ double lds[];
    group_kernel<true, LU / 2>(jobs, tiles, gen, lds);
}

__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6)))
k_group_inv_w6(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles, TileGen gen) {
    extern __shared__ double lds[];
    group_kernel<false, LU / 2>(jobs, tiles, gen, lds);
}

// ---------------------------------------------------------------------------
// Column groups: a strided group of one to three stages with S <= 36 rows
// (radices 7, 5, 3, 4 and their products up to 36).  Its tile is C adjacent
// columns of the S rows; every column is an independent S-point problem of
// the group's stages (the legs of each stage are rows of the same column), so
// thread c loads its column's S points (coalesced across the wave: adjacent
// columns are adjacent in memory), runs every stage of the group on them in
// registers and stores them back.  No LDS and no barrier: the waves only
// stream, and as many stay resident as their registers allow.  The
// butterflies are those of pass1 / pass44, so any schedule gives the same
// spectra.

// One stage of radix M on the column's S points: legs STRIDE rows apart in
// blocks of STRIDE*M rows; the butterfly at row jr of a block has twiddle
// index jr*Lb + q (q = the column), as in pass1.
template <bool FWD, int M, int STRIDE, int S>
__device__ __forceinline__ void col_stage(const Fld &f, double (&y)[S], const StageU &st, int64_t Lb, int64_t q) {
    static_assert(S % (STRIDE * M) == 0, "stage blocks tile the column");
#pragma unroll
    for (int b = 0; b < S / (STRIDE * M); b++)
#pragma unroll
        for (int jr = 0; jr < STRIDE; jr++) {
            const int64_t j = jr * Lb + q;
            double w[M - 1], v[M];
#pragma unroll
            for (int t = 1; t < M; t++) w[t - 1] = gld(st.tw, (t - 1) * st.Q + j);
#pragma unroll
            for (int t = 0; t < M; t++) v[t] = y[b * STRIDE * M + jr + t * STRIDE];
            stage_regs<FWD, M>(f, v, w, st);
#pragma unroll
            for (int t = 0; t < M; t++) y[b * STRIDE * M + jr + t * STRIDE] = v[t];
        }
}

template <bool FWD, int M0, int M1, int M2>
__device__ void col_body(const Fld &f, const DevJob *__restrict__ Jp, const Tile &t) {
    constexpr int S = M0 * (M1 ? M1 : 1) * (M2 ? M2 : 1);
    const auto *J = cst(Jp);
    const DevPlan *__restrict__ P = unip(J->plan[t.prime]);
    const auto *G = &cst(P)->g[t.group];
    const int64_t La = uni64(G->La), Lb = uni64(G->Lb);
    const int logC = uni(G->logC);
    const int s0 = uni(G->s0);
    const int64_t tpr = Lb >> logC;
    const int64_t blk = (int64_t)t.tile / tpr;
    const int64_t jc = (int64_t)t.tile - blk * tpr;
    const int64_t q = (jc << logC) + threadIdx.x;
    const int64_t base = blk * La + q;
    double *__restrict__ x = unip(t.op ? J->xb[t.prime] : J->xa[t.prime]);
    double y[S];
    if ((t.flags & TILE_PACK) != 0) {
        const uint64_t *limbs = unip(t.op ? J->b : J->a);
        const int64_t n = uni64(t.op ? J->n : J->m);
        const int64_t nch = uni64(t.op ? J->ncb : J->nca);
        const int b = uni(J->bch);
        if (n <= BUF_MAX_LIMBS) {
            const __amdgpu_buffer_rsrc_t r = limb_rsrc(limbs, n);
            const uint32_t lmask = b >= 32 ? ~0u : (1u << b) - 1u;
            const uint32_t hmask = b >= 32 ? (1u << (b - 32)) - 1u : 0u;
#pragma unroll
            for (int k = 0; k < S; k++) {
                const int64_t i = base + k * Lb;
                y[k] = pack_buf(r, lmask, hmask, i < nch, i * b);
            }
        } else {
#pragma unroll
            for (int k = 0; k < S; k++) y[k] = pack_chunk(limbs, n, b, nch, base + k * Lb);
        }
    } else {
#pragma unroll
        for (int k = 0; k < S; k++) y[k] = gldn(x, base + k * Lb);
    }
    if (FWD) {
        col_stage<true, M0, S / M0>(f, y, stage_u(&P->st[s0]), Lb, q);
        if constexpr (M1 != 0) col_stage<true, M1, S / (M0 * M1)>(f, y, stage_u(&P->st[s0 + 1]), Lb, q);
        if constexpr (M2 != 0) col_stage<true, M2, 1>(f, y, stage_u(&P->st[s0 + 2]), Lb, q);
    } else {
        if constexpr (M2 != 0) col_stage<false, M2, 1>(f, y, stage_u(&P->st[s0 + 2]), Lb, q);
        if constexpr (M1 != 0) col_stage<false, M1, S / (M0 * M1)>(f, y, stage_u(&P->st[s0 + 1]), Lb, q);
        col_stage<false, M0, S / M0>(f, y, stage_u(&P->st[s0]), Lb, q);
    }
#pragma unroll
    for (int k = 0; k < S; k++) gstn(x, base + k * Lb, y[k]);
}

// The group shapes a column kernel takes (GroupDesc::col = M0 | M1 << 4 | M2 << 8).
template <bool FWD, bool BIG>
__device__ __forceinline__ void col_switch(const Fld &f, const DevJob *J, const Tile &t, int code) {
    if (!BIG) {
        switch (code) {
        case 0x3: col_body<FWD, 3, 0, 0>(f, J, t); break;
        case 0x4: col_body<FWD, 4, 0, 0>(f, J, t); break;
        case 0x5: col_body<FWD, 5, 0, 0>(f, J, t); break;
        case 0x7: col_body<FWD, 7, 0, 0>(f, J, t); break;
        case 0x33: col_body<FWD, 3, 3, 0>(f, J, t); break;
        case 0x43: col_body<FWD, 3, 4, 0>(f, J, t); break;
        case 0x35: col_body<FWD, 5, 3, 0>(f, J, t); break;
        case 0x44: col_body<FWD, 4, 4, 0>(f, J, t); break;
        default: break;
        }
    } else {
        switch (code) {
        case 0x45: col_body<FWD, 5, 4, 0>(f, J, t); break;
        case 0x37: col_body<FWD, 7, 3, 0>(f, J, t); break;
        case 0x47: col_body<FWD, 7, 4, 0>(f, J, t); break;
        case 0x57: col_body<FWD, 7, 5, 0>(f, J, t); break;
        case 0x433: col_body<FWD, 3, 3, 4>(f, J, t); break;
        default: break;
        }
    }
}

template <bool FWD, bool BIG>
__device__ __forceinline__ void col_kernel(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles,
                                           const TileGen &gen) {
    const Tile t = fetch_tile(tiles, gen);
    const DevJob *J = job_of(jobs, t, gen);
    if (uni(cst(J)->zero)) return;
    const DevPlan *P = unip(cst(J)->plan[t.prime ? 1 : 0]);
    const int code = uni(cst(P)->g[t.group].col);
    col_switch<FWD, BIG>(fld(t.prime), J, t, code);
}

// S <= 16: at most 32 VGPRs of points, six waves per SIMD
__global__ void __launch_bounds__(COL_THREADS) __attribute__((amdgpu_waves_per_eu(6)))
k_col_fwd_s(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles, TileGen gen) {
    col_kernel<true, false>(jobs, tiles, gen);
}
__global__ void __launch_bounds__(COL_THREADS) __attribute__((amdgpu_waves_per_eu(6)))
k_col_inv_s(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles, TileGen gen) {
    col_kernel<false, false>(jobs, tiles, gen);
}
// 16 < S <= 36: up to 72 VGPRs of points, three waves per SIMD (four spill)
__global__ void __launch_bounds__(COL_THREADS) __attribute__((amdgpu_waves_per_eu(3)))
k_col_fwd_b(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles, TileGen gen) {
    col_kernel<true, true>(jobs, tiles, gen);
}
__global__ void __launch_bounds__(COL_THREADS) __attribute__((amdgpu_waves_per_eu(3)))
k_col_inv_b(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles, TileGen gen) {
    col_kernel<false, true>(jobs, tiles, gen);
}

// ---------------------------------------------------------------------------
// Mid pass: last forward group on both operands, pointwise product
// (src/fpntt.jl:770-785), first inverse group; for a single-group transform
// the pack is fused in too, so the whole transform pair is one kernel.  A
// job whose b is a cached spectrum (bspec) reads b's finished forward
// spectrum and transforms a only.

__device__ __forceinline__ void mid_body(const Fld &f, const DevJob *__restrict__ Jp, const Tile &t, double *lds) {
    const auto *J = cst(Jp);
    const DevPlan *__restrict__ P = unip(J->plan[t.prime]);
    const auto *G = &cst(P)->g[uni(cst(P)->ngroups) - 1];
    TileGeo g;
    g.S = uni(G->S);
    g.logC = 0;
    g.Lb = 1;
    g.jcol0 = 0;
    const int S = g.S;
    const int s0 = uni(G->s0), s1 = uni(G->s1);
    const int64_t gbase = (int64_t)t.tile * uni64(G->La);
    const bool sq = uni(J->sqr) != 0;
    const bool bspec = uni(J->bspec) != 0;
    double *la = lds;  // the one LDS array
    const bool pack = (t.flags & TILE_PACK) != 0;
    PointSrc sa, sb;
    sa.x = unip(J->xa[t.prime]);
    sb.x = unip(J->xb[t.prime]);
    sa.limbs = unip(J->a);
    sb.limbs = unip(J->b);
    sa.n = uni64(J->m);
    sb.n = uni64(J->n);
    sa.nch = uni64(J->nca);
    sb.nch = uni64(J->ncb);
    sa.b = sb.b = uni(J->bch);
    double hb[HOLD];
    // the first radix-4 pair straight from the loads, the last inverse one
    // straight to HBM (when the pair is not also the last forward pass)
    const bool edge = edge_ok(P, s0, s1, S, 3);
    const int sf = edge ? s0 + 2 : s0;
    if (!sq && !bspec) {
        // b first: its forward passes, then its last pass held in registers
        if (edge) {
            edge_load44<8, true, true>(f, la, sb, pack, g, gbase, P, s0);
        } else {
            load_tile<LU>(la, sb, pack, S, g, gbase);
        }
        __syncthreads();
        const int sl = fwd_until_last(f, la, P, sf, s1, g);
        last_pass<true, SRC_HELD>(f, la, hb, nullptr, 0, P, sl, s1, g);
        __syncthreads();
    }
    if (edge) {
        edge_load44<8, true, true>(f, la, sa, pack, g, gbase, P, s0);
    } else {
        load_tile<LU>(la, sa, pack, S, g, gbase);
    }
    __syncthreads();
    const int sl = fwd_until_last(f, la, P, sf, s1, g);
    if (sq) last_pass<false, SRC_SELF>(f, la, hb, nullptr, 0, P, sl, s1, g);
    else if (bspec) last_pass<false, SRC_GLOB>(f, la, hb, sb.x, gbase, P, sl, s1, g);
    else last_pass<false, SRC_HELD>(f, la, hb, nullptr, 0, P, sl, s1, g);
    __syncthreads();
    run_stages<false, true>(f, la, nullptr, 1, P, sf, sl, g);
    double *__restrict__ xo = unip(J->xa[t.prime]);
    if (edge) edge_store44<true, true>(f, xo, la, g, gbase, P, s0);
    else store_tile(xo, la, S, g, gbase);
}

__global__ void __launch_bounds__(MID_MAX_THREADS) k_mid(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles,
                                              TileGen gen) {
    extern __shared__ double lds[];
    const Tile t = fetch_tile(tiles, gen);
    const DevJob *J = job_of(jobs, t, gen);
    if (uni(cst(J)->zero)) return;
    if ((t.flags & TILE_CLEAR) && t.prime == 0 && t.tile == 0) {
        // the unpack after this launch resolves its carries itself (k_unpack<true>)
        uint32_t *__restrict__ fl = unip(cst(J)->chunk_flags);
        const int64_t nch = uni64(cst(J)->nchunks);
        for (int64_t i = threadIdx.x; i < nch; i += blockDim.x) fl[i] = 0;
    }
    mid_body(fld(t.prime), J, t, lds);
}

// ---------------------------------------------------------------------------
// Garner unpack (src/fpntt.jl:806-925), restated for a parallel machine.
//
// Coefficient i of the product convolution is c_i = c1 + p1*u (< p1*p2 < 2^100)
// with c1, u the reference's canonical Garner pair read at (N-i) mod N and
// scaled by 1/N.  The product is sum_i c_i 2^(b i).  Thread t owns the limbs
// [k0, k0+4): it sums exactly the coefficients that START in those limbs
// (a streamed 192-bit window, as the reference's) into four limb words and a
// two-word spill into the next thread's limbs.  Adding the previous thread's
// spill gives a 4-word value with a carry out; carries between threads are a
// generate/propagate prefix (wave ballot + add trick, then waves in LDS),
// and carries between chunks are resolved by k_carry_fix's look-back.

// The unpack's job constants in scalar registers.
struct UnpackU {
    const double *x1, *x2;
    int64_t N, nconv, rn;
    double n1, n1p, n2, n2p;
    uint64_t bmagic;
    int b;
};

__device__ __forceinline__ UnpackU unpack_u(const DevJob *__restrict__ Jp) {
    UnpackU u;
    const auto *J = cst(Jp);
    const auto *P1 = cst(J->plan[0]);
    const auto *P2 = cst(J->plan[1]);
    u.x1 = unip(J->xa[0]);
    u.x2 = unip(J->xa[1]);
    u.N = uni64(P1->N);
    u.nconv = uni64(J->nconv);
    u.rn = uni64(J->rn);
    u.n1 = unid(P1->ninv);
    u.n1p = unid(P1->ninvp);
    u.n2 = unid(P2->ninv);
    u.n2p = unid(P2->ninvp);
    u.bmagic = (uint64_t)uni64((int64_t)J->bmagic);
    u.b = uni(J->bch);
    return u;
}

// An integer-valued double in [0, 2^52) as an integer: v + 2^52 is exact and
// its mantissa field is v.
__device__ __forceinline__ uint64_t exact_u64(double v) {
    return (uint64_t)__double_as_longlong(v + 4503599627370496.0) - 0x4330000000000000ull;
}

// Garner pair of one coefficient from its two residues (src/fpntt.jl:806-816),
// returned as c = c1 + p1*u < 2^100 in two words.
__device__ __forceinline__ void garner(const UnpackU &U, double x1, double x2, uint64_t &lo, uint64_t &hi) {
    double v1 = fp_mulmod4<0>(x1, U.n1, U.n1p);
    v1 = v1 < 0 ? v1 + P1_D : v1;
    const double v2 = fp_mulmod4<1>(x2, U.n2, U.n2p);
    double u = fp_mulmod4<1>(v2 - fp_reduce<1>(v1), GARNER_G, GARNER_GP);
    u = u < 0 ? u + P2_D : u;
    const unsigned __int128 c = (unsigned __int128)exact_u64(u) * P1_U + exact_u64(v1);
    lo = (uint64_t)c;
    hi = (uint64_t)(c >> 64);
}

struct Seg {
    uint64_t o0, o1, o2, o3;  // limbs k0..k0+3 (before neighbour spill and carries)
    uint64_t s0, s1;          // spill into limbs k0+4, k0+5
};

// Sum of the staged coefficients that start in limbs [k0, k0+4), streamed
// through a 192-bit window as the reference's unpack does.
__device__ __forceinline__ Seg segment_sum(const UnpackU &U, int64_t k0, int64_t i_first,
                                           const uint64_t *clo, const uint64_t *chi) {
    const int b = U.b;
    const int64_t ilo = (int64_t)__umul64hi((uint64_t)(64 * k0 + b - 1), U.bmagic);
    const int64_t iend = (int64_t)__umul64hi((uint64_t)(64 * (k0 + UNPACK_LPT) + b - 1), U.bmagic);
    int s = (int)(ilo * b - 64 * k0);  // in [0, 64) at every add: b < 64, one shift per step
    const int n = (int)(iend - ilo);
    const uint64_t *__restrict__ plo = clo + (ilo - i_first);
    const uint64_t *__restrict__ phi = chi + (ilo - i_first);
    uint64_t w0 = 0, w1 = 0, w2 = 0;
    Seg out{0, 0, 0, 0, 0, 0};
    for (int t = 0; t < n; t++) {
        const uint64_t lo = plo[t], hi = phi[t];
        // (hi:lo) << s in three words; (x >> 1) >> (63 - s) is x >> (64 - s), and 0 for s = 0
        const uint64_t a0 = lo << s;
        const uint64_t a1 = ((lo >> 1) >> (63 - s)) | (hi << s);
        const uint64_t a2 = (hi >> 1) >> (63 - s);
        unsigned long long c0, c1;
        w0 = __builtin_addcll(w0, a0, 0ull, &c0);
        w1 = __builtin_addcll(w1, a1, c0, &c1);
        w2 += a2 + c1;
        s += b;
        if (s >= 64) {
            out.o0 = out.o1; out.o1 = out.o2; out.o2 = out.o3; out.o3 = w0;
            w0 = w1; w1 = w2; w2 = 0;
            s -= 64;
        }
    }
    out.s0 = w0;
    out.s1 = w1;
    return out;
}

// Each chunk first stages the Garner values of every coefficient starting in
// its limbs (and in the 4 limbs before, for the first thread's spill) in LDS:
// coalesced reads of both spectra, each coefficient computed once.  Thread t
// sums the segment of limbs [L0 + 4(t-1), L0 + 4t): thread 0's segment is the
// previous chunk's top one, summed only for the spill it passes to thread 1,
// so every thread sums exactly one segment.
// Held to five waves per SIMD (at most 96 VGPRs): with the staging sized to
// the chunk width, five blocks' LDS fit a CU for b >= 40.
//
// LB (one job whose chunks all fit the GPU at once, UNPACK_LB_MAX_CHUNKS):
// the carries between chunks are resolved here instead of by k_carry_fix.
// Each chunk publishes its summary (carry out with carry-in 0, all-ones) with
// a ready bit as soon as its own sums are done, then walks back over the
// summaries of the chunks below it -- waiting for each to be published -- to
// find its carry-in, and adds it before storing.  A chunk publishes before it
// waits, so the lowest unfinished chunk never waits on anything; the grid is
// far below the blocks one launch can hold resident at once, so every chunk
// waited on is running or gets a slot as other work leaves.  The flags start
// cleared (the mid kernel clears them, TILE_CLEAR) and each wait is bounded.
constexpr uint32_t CHUNK_READY = 4;
constexpr int LB_SPIN_MAX = 1 << 26;

template <bool LB>
__global__ void __launch_bounds__(UNPACK_THREADS) __attribute__((amdgpu_waves_per_eu(5)))
k_unpack(const DevJob *__restrict__ jobs, const Tile *__restrict__ tiles, TileGen gen) {
    extern __shared__ uint64_t ustage[];
    __shared__ uint64_t spill0[UNPACK_THREADS], spill1[UNPACK_THREADS];
    __shared__ uint32_t wsum[UNPACK_THREADS / 64];
    __shared__ int lb_cin, lb_first;
    const Tile t = fetch_tile(tiles, gen);
    const DevJob *__restrict__ Jp = job_of(jobs, t, gen);
    const auto *J = cst(Jp);
    const int tid = threadIdx.x;
    const int64_t L0 = (int64_t)t.tile * UNPACK_CHUNK;
    const int64_t k0 = L0 + UNPACK_LPT * (tid - 1);
    const bool own = tid > 0;  // thread 0's limbs belong to the previous chunk
    uint64_t *__restrict__ r = unip(J->r);
    const int64_t rn = uni64(J->rn);
    uint32_t *__restrict__ flags = unip(J->chunk_flags);
    if (uni(J->zero)) {
        if (own)
            for (int w = 0; w < UNPACK_LPT; w++)
                if (k0 + w < rn) r[k0 + w] = 0;
        if (tid == 0) flags[t.tile] = 0;
        return;
    }
    const UnpackU U = unpack_u(Jp);
    uint64_t *__restrict__ clo = ustage;
    uint64_t *__restrict__ chi = ustage + unpack_stage(U.b);  // the launch's LDS fits its smallest b
    const int64_t kstart = L0 >= UNPACK_LPT ? L0 - UNPACK_LPT : 0;
    const int64_t i_first = (int64_t)__umul64hi((uint64_t)(64 * kstart + U.b - 1), U.bmagic);
    const int64_t i_end = (int64_t)__umul64hi((uint64_t)(64 * (L0 + UNPACK_CHUNK) + U.b - 1), U.bmagic);
    const int cnt = (int)(i_end - i_first);
    for (int l0 = tid; l0 < cnt; l0 += GU * UNPACK_THREADS) {
        double a1[GU], a2[GU];
#pragma unroll
        for (int k = 0; k < GU; k++) {
            const int li = l0 + k * UNPACK_THREADS;
            const int64_t i = i_first + li;
            const bool valid = li < cnt && i < U.nconv;
            const int64_t idx = i == 0 ? 0 : U.N - i;
            a1[k] = valid ? gld(U.x1, idx) : 0.0;
            a2[k] = valid ? gld(U.x2, idx) : 0.0;
        }
#pragma unroll
        for (int k = 0; k < GU; k++) {
            const int li = l0 + k * UNPACK_THREADS;
            if (li < cnt) {
                uint64_t lo = 0, hi = 0;
                if (i_first + li < U.nconv) garner(U, a1[k], a2[k], lo, hi);
                clo[li] = lo;
                chi[li] = hi;
            }
        }
    }
    __syncthreads();
    // chunk 0 has nothing below it: thread 0's segment is empty there
    const Seg sg = k0 >= 0 ? segment_sum(U, k0, i_first, clo, chi) : Seg{0, 0, 0, 0, 0, 0};
    spill0[tid] = sg.s0;
    spill1[tid] = sg.s1;
    __syncthreads();
    uint64_t p0 = 0, p1 = 0;
    if (own) {
        p0 = spill0[tid - 1];
        p1 = spill1[tid - 1];
    }
    // q = o + (p1:p0), four words with a carry out
    uint64_t q0 = sg.o0 + p0;
    uint64_t c = q0 < p0;
    uint64_t q1 = sg.o1 + p1;
    uint64_t c1 = q1 < p1;
    q1 += c;
    c1 += (q1 < c);
    uint64_t q2 = sg.o2 + c1;
    c = q2 < c1;
    uint64_t q3 = sg.o3 + c;
    c = q3 < c;
    // thread 0 neither generates nor stops a carry into the chunk
    const bool G = own && c != 0;
    const bool Pp = !own || (q0 & q1 & q2 & q3) == ~0ull;

    const int lane = tid & 63, wid = tid >> 6;
    const uint64_t gm = __ballot(G);
    const uint64_t pm = __ballot(Pp);
    {
        const uint64_t cv = (gm + (gm | pm)) ^ pm;  // carries into each lane with carry-in 0
        const uint32_t wg = (uint32_t)(((gm >> 63) | ((pm >> 63) & (cv >> 63))) & 1);
        const uint32_t wp = pm == ~0ull ? 1u : 0u;
        if (lane == 0) wsum[wid] = wg | (wp << 1);
    }
    __syncthreads();
    uint64_t cin = 0;
    for (int w = 0; w < wid; w++) {
        const uint32_t f = wsum[w];
        cin = (f & 1) | ((f >> 1) & cin);
    }
    const uint64_t cv = (gm + (gm | pm) + cin) ^ pm;
    const uint64_t my = (cv >> lane) & 1;
    q0 += my;
    c = q0 < my;
    q1 += c;
    c = q1 < c;
    q2 += c;
    c = q2 < c;
    q3 += c;
    if (tid == UNPACK_THREADS - 1) {
        // carry out of the chunk with carry-in 0, and whether all limbs propagate
        const uint32_t cg = (G || (Pp && my)) ? CHUNK_G : 0u;
        uint32_t allp = 1;
        for (int w = 0; w < UNPACK_THREADS / 64; w++) allp &= (wsum[w] >> 1);
        if (LB)
            __hip_atomic_store(&flags[t.tile], cg | (allp ? CHUNK_P : 0u) | CHUNK_READY, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else
            flags[t.tile] = cg | (allp ? CHUNK_P : 0u);
    }
    if (LB) {
        // carry into the chunk = G(ch-1) | (P(ch-1) & carry into ch-1), as k_carry_fix
        if (tid == 0) {
            int cin = 0;
            for (int64_t ch = (int64_t)t.tile - 1; ch >= 0; ch--) {
                uint32_t f = 0;
                for (int spin = 0; spin < LB_SPIN_MAX; spin++) {
                    f = __hip_atomic_load(&flags[ch], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (f & CHUNK_READY) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                if (f & CHUNK_G) {
                    cin = 1;
                    break;
                }
                if (!(f & CHUNK_P)) break;
            }
            lb_cin = cin;
            lb_first = UNPACK_THREADS;
        }
        __syncthreads();
        if (lb_cin) {
            // add it into the leading run of all-ones limbs (thread 0 owns none)
            const bool ones = !own || ((k0 + 0 < rn && q0 == ~0ull) && (k0 + 1 < rn && q1 == ~0ull) &&
                                       (k0 + 2 < rn && q2 == ~0ull) && (k0 + 3 < rn && q3 == ~0ull));
            if (!ones) atomicMin(&lb_first, tid);
            __syncthreads();
            const int first = lb_first;
            if (own && tid < first) {
                q0 = q1 = q2 = q3 = 0;
            } else if (own && tid == first) {
                q0 += 1;
                c = q0 == 0;
                q1 += c;
                c = c && q1 == 0;
                q2 += c;
                c = c && q2 == 0;
                q3 += c;
            }
        }
    }
    if (own) {
        if (k0 + 0 < rn) r[k0 + 0] = q0;
        if (k0 + 1 < rn) r[k0 + 1] = q1;
        if (k0 + 2 < rn) r[k0 + 2] = q2;
        if (k0 + 3 < rn) r[k0 + 3] = q3;
    }
}

// Carry into chunk ch = G(ch-1) | (P(ch-1) & carry into ch-1): one wave per
// chunk (CARRY_FIX_WAVES chunks per block) walks back over the chunk
// summaries (all written by the previous launch, so no waiting), then adds
// the carry into the chunk's leading run of all-ones limbs, 64 limbs a round:
// the run becomes zeros and the first other limb gains one.  Most chunks
// take no carry, so most waves read one summary word and exit.
__global__ void __launch_bounds__(64 * CARRY_FIX_WAVES) k_carry_fix(const DevJob *__restrict__ jobs,
                                                                    const Tile *__restrict__ tiles, TileGen gen,
                                                                    int64_t ntiles) {
    const int lane = threadIdx.x & 63;
    const int64_t idx = (int64_t)blockIdx.x * CARRY_FIX_WAVES + (threadIdx.x >> 6);
    if (idx >= ntiles) return;  // whole waves
    uint32_t job = 0, tile;
    if (tiles) {
        const Tile tt = tiles[idx];
        job = tt.job;
        tile = tt.tile;
    } else {
        tile = (uint32_t)(idx % gen.per);
    }
    const DevJob *__restrict__ J = jobs ? jobs + job : &gen.job;
    if (J->zero || tile == 0) return;
    const uint32_t *__restrict__ flags = J->chunk_flags;
    bool cin = false;
    for (int64_t ch = (int64_t)tile - 1; ch >= 0; ch--) {
        const uint32_t f = flags[ch];
        if (f & CHUNK_G) { cin = true; break; }
        if (!(f & CHUNK_P)) break;
    }
    if (!cin) return;  // the same for the whole wave
    uint64_t *__restrict__ r = J->r;
    const int64_t rn = J->rn;
    const int64_t lo = (int64_t)tile * UNPACK_CHUNK;
    for (int base = 0; base < UNPACK_CHUNK; base += 64) {
        const int64_t k = lo + base + lane;
        const bool in = base + lane < UNPACK_CHUNK && k < rn;
        const uint64_t v = in ? r[k] : 0;
        const uint64_t stop = __ballot(in && v != ~0ull);  // limbs that end the run
        const int first = stop ? __builtin_ctzll(stop) : 64;
        if (in && lane < first) r[k] = 0;
        else if (in && lane == first) r[k] = v + 1;
        if (stop) break;  // a wave-wide value: the loop exits uniformly
    }
}

// ---------------------------------------------------------------------------
// 2x2 cofactor step outputs (the adds and subtractions of hgcd_mul2! and
// hgcd_matrix_adjust!, src/gcd.jl:360-437): one workgroup per output walks
// its limbs in chunks of COMBINE_THREADS * UNPACK_LPT, each thread owning UNPACK_LPT limbs, with
// the carry (borrow) between threads from the same generate/propagate ballot
// prefix as the unpack and between chunks in a register.

// Carry into the calling thread's segment given the block's carry-in; also
// returns the carry out of the whole block.  gen: the segment's carry out with
// carry-in 0; prop: the segment passes a carry-in through.
__device__ __forceinline__ uint32_t block_carry(bool gen, bool prop, uint32_t cin_block, uint32_t *wsum,
                                                uint32_t &cout_block) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const uint64_t gm = __ballot(gen);
    const uint64_t pm = __ballot(prop);
    {
        const uint64_t cv = (gm + (gm | pm)) ^ pm;
        const uint32_t wg = (uint32_t)(((gm >> 63) | ((pm >> 63) & (cv >> 63))) & 1);
        const uint32_t wp = pm == ~0ull ? 1u : 0u;
        if (lane == 0) wsum[wid] = wg | (wp << 1);
    }
    __syncthreads();
    uint64_t cin = cin_block;
    for (int w = 0; w < wid; w++) {
        const uint32_t f = wsum[w];
        cin = (f & 1) | ((f >> 1) & cin);
    }
    uint32_t co = cin_block;
    for (int w = 0; w < nw; w++) {
        const uint32_t f = wsum[w];
        co = (f & 1) | ((f >> 1) & co);
    }
    cout_block = co;
    const uint64_t cv = (gm + (gm | pm) + cin) ^ pm;
    __syncthreads();  // wsum is reused by the next call
    return (uint32_t)((cv >> lane) & 1);
}

// ---------------------------------------------------------------------------
// A whole product in one workgroup, for single products of short transforms
// (the launcher admits N when the arrays below fit one block's LDS): the chunk
// pack (src/fpntt.jl:744-768), both primes' forward transforms of both
// operands (:703-733), the pointwise products (:770-785), the inverse
// transforms, the Garner CRT and the limb sums with their carry chain
// (:806-925) all run on LDS, so only operand limbs are read and result limbs
// written, by one launch.  Both primes' twiddle tables are copied to LDS in
// the same round of loads as the limbs, so no pass waits on a twiddle load.
// The lower half of the block works on p1, the upper half on p2 (each half a
// whole number of waves, so the split is wave-uniform); the radices of a
// length depend on N only, so both halves run the same pass sequence and
// every barrier is reached by every wave.  The butterflies are those of pass44
// and pass1 (the same stages, the same reduce/mulmod order).
//
// LDS (doubles, each array padded as pidx): a's points for p1 and p2, b's for
// p1 and p2 (ld each), then the two twiddle tables (twl each).  After the
// inverse transforms, b's arrays hold the staged Garner words and a's the
// spills between threads.

// One stage's constants for a whole-transform pass (rows = N, one column):
// the row stride q is the stage's Q, and tw points at its LDS twiddle rows.
// sd is the kernel's LDS copy of the stage descriptor: a scalar load from the
// plan would be an L2 round trip at every pass of so short a kernel.
__device__ __forceinline__ StageU small_stage(const StageDesc *sd, const double *twl) {
    StageU u;
    u.m = uni(sd->m);
    u.Q = uni64(sd->Q);
    u.q = (int)u.Q;
    u.qmagic = (uint64_t)(0xFFFFFFFFu / (uint32_t)u.q) + 1;  // ceil(2^32 / q)
    u.tw = twl + uni64(sd->tw_off);
#pragma unroll
    for (int i = 0; i < 8; i++) u.rot[i] = unid(sd->rot[i]);
    return u;
}

// Radix-4 pair (stages s0, s1) over narr arrays, by the lt-th of H threads.
template <bool FWD, bool LIN>
__device__ void small_pass44(const Fld &f, double *__restrict__ x0, double *__restrict__ x1, int narr, const StageU &s0,
                             const StageU &s1, int N, int lt, int H) {
    TileGeo g;
    g.S = N;
    g.logC = 0;
    g.Lb = 1;
    g.jcol0 = 0;
    const int nsub = N >> 4;
    for (int sg = lt; sg < nsub * narr; sg += H) {
        double *__restrict__ x = sg < nsub ? x0 : x1;
        const Sub44 u = sub44(sg < nsub ? sg : sg - nsub, s0, s1, g);
        double w0[4][3], w1[3];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int t = 0; t < 3; t++) w0[a][t] = s0.tw[t * s0.Q + u.jt + a * s1.Q];
#pragma unroll
        for (int t = 0; t < 3; t++) w1[t] = s1.tw[t * s1.Q + u.jt];
        double y[4][4];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) y[a][b] = x[ix44<LIN>(u, a, b)];
        regs44<FWD>(f, y, w0, w1, s0, s1);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) x[ix44<LIN>(u, a, b)] = y[a][b];
    }
}

// One radix-M stage over narr arrays, by the lt-th of H threads.
template <bool FWD, int M>
__device__ void small_pass1(const Fld &f, double *__restrict__ x0, double *__restrict__ x1, int narr, const StageU &s, int N,
                            int lt, int H) {
    TileGeo g;
    g.S = N;
    g.logC = 0;
    g.Lb = 1;
    g.jcol0 = 0;
    const int nb = N / M;
    for (int bg = lt; bg < nb * narr; bg += H) {
        double *__restrict__ x = bg < nb ? x0 : x1;
        const Sub1 u = sub1<M>(bg < nb ? bg : bg - nb, s, g);
        double w[M > 1 ? M - 1 : 1] = {};
        if constexpr (M > 2) {
#pragma unroll
            for (int t = 1; t < M; t++) w[t - 1] = s.tw[(t - 1) * s.Q + u.j];
        }
        double y[M];
#pragma unroll
        for (int t = 0; t < M; t++) y[t] = x[ix1<false>(u, t)];
        stage_regs<FWD, M>(f, y, w, s);
#pragma unroll
        for (int t = 0; t < M; t++) x[ix1<false>(u, t)] = y[t];
    }
}

// Stages [s, s + (pair ? 2 : 1)) of prime K's plan over its narr arrays.
template <bool FWD>
__device__ __forceinline__ void small_pass(const Fld &f, double *x0, double *x1, int narr, const StageDesc *st, int s, bool pair,
                                           const double *twl, int N, int lt, int H) {
    const StageU a = small_stage(&st[s], twl);
    if (pair) {
        const StageU b = small_stage(&st[s + 1], twl);
        if (((a.q | b.q) & 15) == 0) small_pass44<FWD, true>(f, x0, x1, narr, a, b, N, lt, H);
        else small_pass44<FWD, false>(f, x0, x1, narr, a, b, N, lt, H);
        return;
    }
    switch (a.m) {
    case 4: small_pass1<FWD, 4>(f, x0, x1, narr, a, N, lt, H); break;
    case 2: small_pass1<FWD, 2>(f, x0, x1, narr, a, N, lt, H); break;
    case 3: small_pass1<FWD, 3>(f, x0, x1, narr, a, N, lt, H); break;
    case 5: small_pass1<FWD, 5>(f, x0, x1, narr, a, N, lt, H); break;
    default: small_pass1<FWD, 7>(f, x0, x1, narr, a, N, lt, H); break;
    }
}

// Loads per thread per round of the small kernel's first copies: chunks of
// a and b, both twiddle tables (the launcher gives the block at least a
// quarter as many threads as points, and the tables hold at most about
// 1.1 N doubles, so one round covers them).
constexpr int SMALL_LU = 4;
// Doubles of one stage descriptor (the kernel copies both plans' to LDS).
constexpr int STAGE_WORDS = (int)(sizeof(StageDesc) / sizeof(double));
static_assert(sizeof(StageDesc) % sizeof(double) == 0, "stage descriptors copy as doubles");

// single: run every stage as its own pass (one radix-4 butterfly per thread
// and array when the block has a thread per 4 points of a prime's arrays:
// four times the waves of the 16-point register pairs, which a block this
// small needs to hide its FP64 and LDS latencies); otherwise radix-4 pairs.
__global__ void __launch_bounds__(SMALL_MAX_THREADS) k_small(const DevJob J, int ld, int twl, int nseg, int cnt,
                                                             int single) {
    extern __shared__ double lds[];
    __shared__ StageDesc sst[2][SMALL_MAX_STAGES];
    const int tid = threadIdx.x, NT = blockDim.x, H = NT >> 1;
    const int k = uni(tid >= H ? 1 : 0);  // prime of this thread's half (H is a multiple of 64)
    const Fld fk = fld(k);
    const int lt = tid - k * H;
    const DevPlan *__restrict__ P0 = unip(J.plan[0]);
    const DevPlan *__restrict__ P1 = unip(J.plan[1]);
    const int N = (int)uni64(cst(P0)->N);
    const int nst = uni(cst(P0)->nst);
    const bool sq = J.sqr != 0;
    const int narr = sq ? 1 : 2;
    double *const xa0 = lds, *const xa1 = lds + ld, *const xb0 = lds + 2 * ld, *const xb1 = lds + 3 * ld;
    double *const tw0 = lds + 4 * ld, *const tw1 = tw0 + twl;
    double *const xa = k ? xa1 : xa0, *const xb = k ? xb1 : xb0;
    const double *const twk = k ? tw1 : tw0;
    const StageDesc *const stk = sst[k];

    // chunks of a (and b) for both primes, both twiddle tables and both
    // plans' stage descriptors: every load of a round is issued before any
    // result is used, so the copies cost one memory latency
    {
        const int b = J.bch;
        const uint32_t lmask = b >= 32 ? ~0u : (1u << b) - 1u;
        const uint32_t hmask = b >= 32 ? (1u << (b - 32)) - 1u : 0u;
        const __amdgpu_buffer_rsrc_t ra = limb_rsrc(J.a, J.m), rb = limb_rsrc(J.b, J.n);
        const double *__restrict__ g0 = unip(cst(P0)->twc);
        const double *__restrict__ g1 = unip(cst(P1)->twc);
        const int sw = STAGE_WORDS * nst;
        const int top = max(max(N, twl), 2 * sw);
        for (int r0 = 0; r0 < top; r0 += SMALL_LU * NT) {
            double va[SMALL_LU], vb[SMALL_LU], v0[SMALL_LU], v1[SMALL_LU];
            const int i0 = r0 + tid;
            const int pk = i0 >= sw ? 1 : 0, w = i0 - pk * sw;
            const double *sd = (const double *)(pk ? &P1->st[0] : &P0->st[0]);
            const double vs = gld(sd, i0 < 2 * sw ? w : 0);
#pragma unroll
            for (int u = 0; u < SMALL_LU; u++) {
                const int i = r0 + u * NT + tid;
                va[u] = pack_buf(ra, lmask, hmask, i < J.nca, (int64_t)i * b);
                vb[u] = sq ? 0.0 : pack_buf(rb, lmask, hmask, i < J.ncb, (int64_t)i * b);
                const int it = min(i, twl - 1);
                v0[u] = gld(g0, it);
                v1[u] = gld(g1, it);
            }
            if (i0 < 2 * sw) ((double *)&sst[pk][0])[w] = vs;
#pragma unroll
            for (int u = 0; u < SMALL_LU; u++) {
                const int i = r0 + u * NT + tid;
                if (i < N) {
                    const int p = pidx(i);
                    xa0[p] = va[u];
                    xa1[p] = va[u];
                    if (!sq) {
                        xb0[p] = vb[u];
                        xb1[p] = vb[u];
                    }
                }
                if (i < twl) {
                    tw0[i] = v0[u];
                    tw1[i] = v1[u];
                }
            }
        }
    }
    __syncthreads();

    // forward transforms of a and b (radix-4 pairs from the front)
    for (int s = 0; s < nst;) {
        const bool pair = !single && uni(sst[0][s].m) == 4 && s + 1 < nst && uni(sst[0][s + 1].m) == 4;
        small_pass<true>(fk, xa, xb, narr, stk, s, pair, twk, N, lt, H);
        __syncthreads();
        s += pair ? 2 : 1;
    }
    // pointwise product (a square multiplies a by itself)
    {
        const double *__restrict__ yb = sq ? xa : xb;
        for (int i = lt; i < N; i += H) {
            const int p = pidx(i);
            xa[p] = fp_mulmod3(fk, xa[p], yb[p]);
        }
    }
    __syncthreads();
    // inverse transforms of the products (radix-4 pairs from the back)
    for (int s = nst - 1; s >= 0;) {
        const bool pair = !single && uni(sst[0][s].m) == 4 && s - 1 >= 0 && uni(sst[0][s - 1].m) == 4;
        const int s0 = pair ? s - 1 : s;
        small_pass<false>(fk, xa, nullptr, 1, stk, s0, pair, twk, N, lt, H);
        __syncthreads();
        s -= pair ? 2 : 1;
    }

    // Garner words of coefficients [0, cnt) (zero past nconv) into b's arrays
    UnpackU U;
    U.x1 = U.x2 = nullptr;
    U.N = N;
    U.nconv = J.nconv;
    U.rn = J.rn;
    U.n1 = unid(cst(P0)->ninv);
    U.n1p = unid(cst(P0)->ninvp);
    U.n2 = unid(cst(P1)->ninv);
    U.n2p = unid(cst(P1)->ninvp);
    U.bmagic = J.bmagic;
    U.b = J.bch;
    uint64_t *const clo = (uint64_t *)xb0, *const chi = (uint64_t *)xb1;
    for (int i = tid; i < cnt; i += NT) {
        uint64_t lo = 0, hi = 0;
        if (i < U.nconv) {
            const int p = pidx(i == 0 ? 0 : N - i);
            garner(U, xa0[p], xa1[p], lo, hi);
        }
        clo[i] = lo;
        chi[i] = hi;
    }
    __syncthreads();
    // thread t sums the coefficients starting in limbs [4t, 4t + 4) (one round:
    // the launcher makes the block at least nseg threads)
    uint64_t *const spill0 = (uint64_t *)xa0, *const spill1 = spill0 + nseg;
    uint32_t *const wsum = (uint32_t *)xa1;
    const int64_t k0 = (int64_t)UNPACK_LPT * tid;
    const bool mine = tid < nseg;
    const Seg sg = mine ? segment_sum(U, k0, 0, clo, chi) : Seg{0, 0, 0, 0, 0, 0};
    if (mine) {
        spill0[tid] = sg.s0;
        spill1[tid] = sg.s1;
    }
    __syncthreads();
    const bool below = tid > 0 && tid <= nseg;  // threads with a segment under them
    const uint64_t p0 = below ? spill0[tid - 1] : 0, p1 = below ? spill1[tid - 1] : 0;
    uint64_t q0 = sg.o0 + p0;
    uint64_t c = q0 < p0;
    uint64_t q1 = sg.o1 + p1;
    uint64_t c1 = q1 < p1;
    q1 += c;
    c1 += (q1 < c);
    uint64_t q2 = sg.o2 + c1;
    c = q2 < c1;
    uint64_t q3 = sg.o3 + c;
    c = q3 < c;
    uint32_t cout = 0;
    const uint32_t my = block_carry(c != 0, (q0 & q1 & q2 & q3) == ~0ull, 0, wsum, cout);
    q0 += my;
    c = q0 < my;
    q1 += c;
    c = q1 < c;
    q2 += c;
    c = q2 < c;
    q3 += c;
    if (mine) {
        uint64_t *__restrict__ r = J.r;
        const int64_t rn = J.rn;
        if (k0 + 0 < rn) r[k0 + 0] = q0;
        if (k0 + 1 < rn) r[k0 + 1] = q1;
        if (k0 + 2 < rn) r[k0 + 2] = q2;
        if (k0 + 3 < rn) r[k0 + 3] = q3;
    }
}

// r[0..lr) = X + Y*beta^ysh (add) or X - Y*beta^ysh (sub, X >= Y*beta^ysh):
// one block walks the limbs in chunks of blockDim * UNPACK_LPT, the carry (borrow)
// between threads from the ballot prefix and between chunks in a register.
// A thread reads limbs k0..k0+3 of X and Y before it writes those of r, so r
// may be X or Y.
// Returns 1 + the highest index k < lr with r[k] != 0 among the calling
// thread's limbs (0 if none).
__device__ __forceinline__ int64_t combine_pass(const uint64_t *X, int64_t lx, const uint64_t *Y, int64_t ly,
                                                int64_t ysh, uint64_t *r, int64_t lr, bool sub, uint32_t *wsum) {
    const int tid = threadIdx.x;
    const int lpc = (int)blockDim.x * UNPACK_LPT;  // limbs per block-wide step
    uint32_t carry = 0;
    int64_t top = 0;
    for (int64_t c0 = 0; c0 < lr; c0 += lpc) {
        const int64_t k0 = c0 + UNPACK_LPT * tid;
        uint64_t d[UNPACK_LPT];
        bool prop = true;
        // segment result with carry (borrow) in 0, and its generate/propagate
        uint64_t cc = 0;
#pragma unroll
        for (int w = 0; w < UNPACK_LPT; w++) {
            const int64_t k = k0 + w;
            const uint64_t x = k < lx ? gldu(X, k) : 0;
            const uint64_t y = k >= ysh && k - ysh < ly ? gldu(Y, k - ysh) : 0;
            if (!sub) {
                const uint64_t s1 = x + y;
                const uint64_t c1 = s1 < x;
                const uint64_t s2 = s1 + cc;
                cc = c1 | (s2 < s1);
                d[w] = s2;
                prop = prop && s2 == ~0ull;
            } else {
                const uint64_t s1 = x - y;
                const uint64_t b1 = x < y;
                const uint64_t s2 = s1 - cc;
                cc = b1 | (s1 < cc);
                d[w] = s2;
                prop = prop && s2 == 0;
            }
        }
        uint32_t cout = 0;
        const uint32_t cin = block_carry(cc != 0, prop, carry, wsum, cout);
        uint64_t cv = cin;
#pragma unroll
        for (int w = 0; w < UNPACK_LPT; w++) {
            const int64_t k = k0 + w;
            uint64_t v;
            if (!sub) {
                v = d[w] + cv;
                cv = v < cv;
            } else {
                v = d[w] - cv;
                cv = d[w] < cv;
            }
            if (k < lr) {
                r[k] = v;
                if (v) top = k + 1;
            }
        }
        carry = cout;
    }
    return top;
}

__global__ void __launch_bounds__(COMBINE_THREADS) k_combine(const CombineJob *__restrict__ jobs) {
    __shared__ uint32_t wsum[COMBINE_THREADS / 64];
    __shared__ unsigned long long top;  // 1 + highest differing limb, 0 if none
    const auto *J = cst(jobs + blockIdx.x);
    const uint64_t *A = unip(J->A), *B = unip(J->B);
    int64_t la = uni64(J->la);
    const int64_t lb = uni64(J->lb), lr = uni64(J->lr);
    uint64_t *__restrict__ r = unip(J->r);
    const bool sub = uni(J->sub) != 0;
    const int tid = threadIdx.x;
    const uint64_t *C = unip(J->C);
    if (C) {
        // hgcd_matrix_adjust!'s shifted top (src/gcd.jl:415-419): A + C*beta^cshift into r first
        combine_pass(A, la, C, uni64(J->lc), uni64(J->cshift), r, lr, false, wsum);
        __syncthreads();  // r, written by the whole block, is read below by other threads
        A = r;
        la = lr;
    }
    bool swap = false;
    if (sub) {
        // highest limb where A and B differ decides the sign: scanned from the
        // top, a block-width of limbs at a time, so it stops at the first
        // window that differs (almost always the top one) instead of reading
        // both operands whole
        if (tid == 0) top = 0;
        __syncthreads();
        const int64_t L = la > lb ? la : lb;
        for (int64_t hi0 = L; hi0 > 0; hi0 -= blockDim.x) {
            const int64_t k = hi0 - 1 - tid;
            if (k >= 0) {
                const uint64_t a = k < la ? gldu(A, k) : 0, b = k < lb ? gldu(B, k) : 0;
                if (a != b) atomicMax(&top, (unsigned long long)k + 1);
            }
            __syncthreads();
            const bool found = top != 0;
            __syncthreads();  // every thread has read top before the next window may raise it
            if (found) break;  // the same value in every thread: the loop exits uniformly
        }
        const int64_t t = (int64_t)top - 1;
        int32_t sg = 0;
        if (t >= 0) {
            const uint64_t a = t < la ? gldu(A, t) : 0, b = t < lb ? gldu(B, t) : 0;
            sg = a > b ? 1 : -1;
        }
        if (tid == 0) *unip(J->sign) = sg;
        swap = sg < 0;
        __syncthreads();  // every thread has read A before r (which may be A) is rewritten
    }
    const uint64_t *X = swap ? B : A, *Y = swap ? A : B;
    const int64_t lx = swap ? lb : la, ly = swap ? la : lb;
    const int64_t mytop = combine_pass(X, lx, Y, ly, 0, r, lr, sub, wsum);
    int64_t *len = unip(J->len);
    if (len) {
        // normlen of r: block maximum, then one atomic on the shared length
        if (tid == 0) top = 0;
        __syncthreads();
        if (mytop) atomicMax(&top, (unsigned long long)mytop);
        __syncthreads();
        if (tid == 0 && top) atomicMax((unsigned long long *)len, top);
    }
}

// Sets every combine job's length output to 1 (normlen's floor) before the
// combine launch takes maxima into it.
__global__ void __launch_bounds__(256) k_combine_init(const CombineJob *__restrict__ jobs, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t *len = jobs[i].len;
    if (len) *len = 1;
}

// One workgroup per reduction.  Thread t owns limbs c0 + 4t .. c0 + 4t + 3 of
// every 1024-limb chunk c0 in all three passes, so each pass reads back only
// limbs the same thread wrote.
__device__ __forceinline__ void barrett_sub_pass(const uint64_t *__restrict__ X, const uint64_t *__restrict__ Y,
                                                 int64_t ly, uint64_t *__restrict__ w, int64_t L, uint32_t *wsum) {
    uint32_t borrow = 0;
    for (int64_t c0 = 0; c0 < L; c0 += UNPACK_LPC) {
        const int64_t k0 = c0 + UNPACK_LPT * threadIdx.x;
        uint64_t d[UNPACK_LPT];
        bool prop = true;
        uint64_t cc = 0;
#pragma unroll
        for (int q = 0; q < UNPACK_LPT; q++) {
            const int64_t k = k0 + q;
            const uint64_t x = k < L ? X[k] : 0, y = k < ly ? Y[k] : 0;
            const uint64_t s1 = x - y;
            const uint64_t b1 = x < y;
            const uint64_t s2 = s1 - cc;
            cc = b1 | (s1 < cc);
            d[q] = s2;
            prop = prop && s2 == 0;
        }
        uint32_t cout = 0;
        const uint32_t cin = block_carry(cc != 0, prop, borrow, wsum, cout);
        uint64_t cv = cin;
#pragma unroll
        for (int q = 0; q < UNPACK_LPT; q++) {
            const int64_t k = k0 + q;
            const uint64_t v = d[q] - cv;
            cv = d[q] < cv;
            if (k < L) w[k] = v;
        }
        borrow = cout;
    }
}

__global__ void __launch_bounds__(UNPACK_THREADS) k_barrett_fix(const BarrettFixJob *__restrict__ jobs) {
    __shared__ uint32_t wsum[UNPACK_THREADS / 64];
    __shared__ unsigned long long top;  // 1 + highest limb where w and m differ, 0 if equal
    __shared__ int ge;
    const auto *J = cst(jobs + blockIdx.x);
    const uint64_t *T = unip(J->T), *P = unip(J->P), *m = unip(J->m);
    uint64_t *__restrict__ w = unip(J->w);
    uint64_t *__restrict__ r = unip(J->r);
    const int64_t k = uni64(J->k), L = k + 1;
    barrett_sub_pass(T, P, L, w, L, wsum);
    for (int rep = 0; rep < 2; rep++) {
        if (threadIdx.x == 0) { top = 0; ge = 0; }
        __syncthreads();
        unsigned long long hi = 0;
        for (int64_t c0 = 0; c0 < L; c0 += UNPACK_LPC)
#pragma unroll
            for (int q = 0; q < UNPACK_LPT; q++) {
                const int64_t i = c0 + UNPACK_LPT * threadIdx.x + q;
                if (i < L && w[i] != (i < k ? m[i] : 0)) hi = (unsigned long long)i + 1;
            }
        if (hi) atomicMax(&top, hi);
        __syncthreads();
        const int64_t t = (int64_t)top - 1;
        // the owner of limb t decides w >= m; equal values count as >=
        if (t < 0) {
            if (threadIdx.x == 0) ge = 1;
        } else if (((t % UNPACK_LPC) >> 2) == threadIdx.x) {
            ge = w[t] > (t < k ? m[t] : 0) ? 1 : 0;
        }
        __syncthreads();
        if (!ge) break;
        barrett_sub_pass(w, m, k, w, L, wsum);
    }
    for (int64_t c0 = 0; c0 < L; c0 += UNPACK_LPC)
#pragma unroll
        for (int q = 0; q < UNPACK_LPT; q++) {
            const int64_t i = c0 + UNPACK_LPT * threadIdx.x + q;
            if (i < k) r[i] = w[i];
        }
}

// ---------------------------------------------------------------------------
// Schoolbook products below the NTT thresholds (the results of mul_basecase! /
// sqr_basecase! / Karatsuba, src/kernels/mul.jl:305-320,406, src/mul.jl:
// 66-118: the product is unique, so any exact method gives the reference's
// limbs).  One workgroup per product.  Thread u owns result limbs
// [c0 + 4u, c0 + 4u + 4) of each 1024-limb chunk: it sums the four columns
// S_t = sum_j a[t-j] b[j] (< 2^136, three words each) that start in those
// limbs into six words -- four of its own and a two-word spill into the
// next thread's limbs.  Adding the previous thread's spill leaves at most one
// carry out of the four words; carries between threads are the block
// generate/propagate prefix, between chunks a register (as in the unpack).

__device__ __forceinline__ void add3(uint64_t &c0, uint64_t &c1, uint64_t &c2, uint64_t lo, uint64_t hi) {
    c0 += lo;
    const uint64_t k0 = c0 < lo;
    const uint64_t t = hi + k0;  // hi < 2^64 - 1, so no overflow here
    c1 += t;
    c2 += c1 < t;
}

__device__ void basecase_body(const BaseJob &J, uint64_t *sb, uint64_t (*spill)[2], uint32_t *wsum) {
    const uint64_t *__restrict__ a = J.a;
    const int64_t m = J.m, n = J.n, rn = J.rn, top = m + n - 1;  // columns >= top hold no products
    for (int j = threadIdx.x; j < n; j += blockDim.x) sb[j] = J.b[j];
    if (threadIdx.x == 0) spill[0][0] = spill[0][1] = 0;
    __syncthreads();
    uint32_t carry = 0;
    for (int64_t c0 = 0; c0 < rn; c0 += UNPACK_LPC) {
        const int64_t k0 = c0 + UNPACK_LPT * threadIdx.x;
        uint64_t W[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < UNPACK_LPT; i++) {
            const int64_t t = k0 + i;
            uint64_t x0 = 0, x1 = 0, x2 = 0;
            if (t < top) {
                const int64_t jlo = t - m + 1 > 0 ? t - m + 1 : 0;
                const int64_t jhi = t < n - 1 ? t : n - 1;
                for (int64_t j = jlo; j <= jhi; j++) {
                    const uint64_t x = gldu(a, t - j), y = sb[j];
                    add3(x0, x1, x2, x * y, __umul64hi(x, y));
                }
            }
            // W += (x0, x1, x2) << 64 i
            uint64_t cc = 0;
#pragma unroll
            for (int w = 0; w < 6; w++) {
                const uint64_t add = w == i ? x0 : w == i + 1 ? x1 : w == i + 2 ? x2 : 0;
                const uint64_t s1 = W[w] + add;
                const uint64_t k1 = s1 < add;
                const uint64_t s2 = s1 + cc;
                cc = k1 | (s2 < s1);
                W[w] = s2;
            }
        }
        spill[threadIdx.x + 1][0] = W[4];
        spill[threadIdx.x + 1][1] = W[5];
        __syncthreads();
        const uint64_t p0 = spill[threadIdx.x][0], p1 = spill[threadIdx.x][1];
        // own four words + the previous thread's spill: one carry out at most
        uint64_t o[4];
        uint64_t cc = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const uint64_t add = w == 0 ? p0 : w == 1 ? p1 : 0;
            const uint64_t s1 = W[w] + add;
            const uint64_t k1 = s1 < add;
            const uint64_t s2 = s1 + cc;
            cc = k1 | (s2 < s1);
            o[w] = s2;
        }
        const bool prop = (o[0] & o[1] & o[2] & o[3]) == ~0ull;
        __syncthreads();  // every thread has read its spill slot
        if (threadIdx.x == blockDim.x - 1) {
            spill[0][0] = W[4];
            spill[0][1] = W[5];
        }
        uint32_t cout = 0;
        const uint32_t cin = block_carry(cc != 0, prop, carry, wsum, cout);
        uint64_t cv = cin;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const int64_t k = k0 + w;
            const uint64_t v = o[w] + cv;
            cv = v < cv;
            if (k < rn) J.r[k] = v;
        }
        carry = cout;
    }
}

__global__ void __launch_bounds__(UNPACK_THREADS) k_basecase(const BaseJob *__restrict__ jobs) {
    __shared__ uint64_t sb[BASE_NMAX];
    __shared__ uint64_t spill[UNPACK_THREADS + 1][2];
    __shared__ uint32_t wsum[UNPACK_THREADS / 64];
    const auto *Jc = cst(jobs + blockIdx.x);
    BaseJob J;
    J.a = unip(Jc->a);
    J.b = unip(Jc->b);
    J.r = unip(Jc->r);
    J.m = uni64(Jc->m);
    J.n = uni64(Jc->n);
    J.rn = uni64(Jc->rn);
    basecase_body(J, sb, spill, wsum);
}

// One schoolbook product passed by value (single calls need no job table).
__global__ void __launch_bounds__(UNPACK_THREADS) k_basecase_one(BaseJob J) {
    __shared__ uint64_t sb[BASE_NMAX];
    __shared__ uint64_t spill[UNPACK_THREADS + 1][2];
    __shared__ uint32_t wsum[UNPACK_THREADS / 64];
    basecase_body(J, sb, spill, wsum);
}

// ---------------------------------------------------------------------------
// Twiddle tables.  Row t (1..m-1) of a stage of span L holds w_L^(t j) for
// j < Q, w_L = w_N^(N/L), minimal balanced (src/fpntt.jl:173-225).  Each entry
// is w_N^e = hi[e >> h] * lo[e & (2^h-1)] mod p with exact host-built tables;
// the mulmod is exact and fp_minbal picks the unique |w| <= p/2 representative,
// so the table equals the reference's chained fp_twiddles output bit for bit.

template <int K>
__global__ void k_twiddle_stage(double *__restrict__ out, int m, int64_t Q, int64_t NoverL,
                                const double *__restrict__ tlo, const double *__restrict__ thi, int h) {
    const int64_t total = (int64_t)(m - 1) * Q;
    const int64_t mask = (1ll << h) - 1;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = idx / Q + 1;
        const int64_t j = idx - (t - 1) * Q;
        const int64_t e = t * j * NoverL;
        out[idx] = fp_minbal<K>(fp_mulmod3<K>(thi[e >> h], tlo[e & mask]));
    }
}

template __global__ void k_twiddle_stage<0>(double *, int, int64_t, int64_t, const double *, const double *, int);
template __global__ void k_twiddle_stage<1>(double *, int, int64_t, int64_t, const double *, const double *, int);

// ---------------------------------------------------------------------------
// Stage-level kernels behind the debug/parity entry points.

template <int K>
__global__ void k_pointwise(double *__restrict__ xa, const double *__restrict__ xb, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        xa[i] = fp_mulmod3<K>(xa[i], xb[i]);
}
template __global__ void k_pointwise<0>(double *, const double *, int64_t);
template __global__ void k_pointwise<1>(double *, const double *, int64_t);

__global__ void k_pack(double *__restrict__ x, int64_t N, const uint64_t *__restrict__ limbs, int64_t n, int b,
                       int64_t nch) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = pack_chunk(limbs, n, b, nch, i);
}

__global__ void k_fill_u64(uint64_t *__restrict__ p, int64_t n, uint64_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Expand a batch's tile ranges into its tile list (one block per range).
__global__ void k_expand_tiles(const TileRange *__restrict__ ranges, Tile *__restrict__ tiles) {
    const TileRange rg = ranges[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < rg.count; i += blockDim.x) {
        Tile t = rg.head;
        t.tile = rg.head.tile + i;
        tiles[rg.out + i] = t;
    }
}

// Top limbs of every operand of a batch (magnitude_bits needs them,
// src/kernels/addsub.jl:150).
__global__ void k_top_limbs(const DevJob *__restrict__ jobs, int64_t count, uint64_t *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        out[2 * i] = jobs[i].a[jobs[i].m - 1];
        out[2 * i + 1] = jobs[i].sqr ? jobs[i].a[jobs[i].m - 1] : jobs[i].b[jobs[i].n - 1];
    }
}

}  // namespace nbh

// ---------------------------------------------------------------------------
// Host launchers (nbh_launch.h)

#include "nbh_launch.h"

namespace nbh {

static inline unsigned grid_for(int64_t n, int threads, int64_t cap = 1 << 16) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

hipError_t launch_group(const DevJob *jobs, const Tile *tiles, const TileGen &gen, int64_t ntiles, int fwd,
                        size_t lds_bytes, int threads, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    if (lds_bytes == COL_SMALL || lds_bytes == COL_BIG) {
        const bool big = lds_bytes == COL_BIG;
        const dim3 grid((unsigned)ntiles), block(threads);
        if (fwd) {
            if (big) hipLaunchKernelGGL(k_col_fwd_b, grid, block, 0, s, jobs, tiles, gen);
            else hipLaunchKernelGGL(k_col_fwd_s, grid, block, 0, s, jobs, tiles, gen);
        } else {
            if (big) hipLaunchKernelGGL(k_col_inv_b, grid, block, 0, s, jobs, tiles, gen);
            else hipLaunchKernelGGL(k_col_inv_s, grid, block, 0, s, jobs, tiles, gen);
        }
        return hipGetLastError();
    }
    // waves per CU that LDS admits: above 16 (four per SIMD) the six-wave build pays
    const bool w6 = group_six_waves(lds_bytes, threads);
    const dim3 grid((unsigned)ntiles), block(threads);
    if (fwd) {
        if (w6) hipLaunchKernelGGL(k_group_fwd_w6, grid, block, lds_bytes, s, jobs, tiles, gen);
        else hipLaunchKernelGGL(k_group_fwd, grid, block, lds_bytes, s, jobs, tiles, gen);
    } else {
        if (w6) hipLaunchKernelGGL(k_group_inv_w6, grid, block, lds_bytes, s, jobs, tiles, gen);
        else hipLaunchKernelGGL(k_group_inv, grid, block, lds_bytes, s, jobs, tiles, gen);
    }
    return hipGetLastError();
}

hipError_t launch_mid(const DevJob *jobs, const Tile *tiles, const TileGen &gen, int64_t ntiles, size_t lds_bytes,
                      int threads, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mid, dim3((unsigned)ntiles), dim3(threads), lds_bytes, s, jobs, tiles, gen);
    return hipGetLastError();
}

hipError_t launch_small(const DevJob &J, int threads, int ld, int twl, int nseg, int cnt, int single,
                        size_t lds_bytes, hipStream_t s) {
    hipLaunchKernelGGL(k_small, dim3(1), dim3(threads), lds_bytes, s, J, ld, twl, nseg, cnt, single);
    return hipGetLastError();
}

hipError_t launch_unpack(const DevJob *jobs, const Tile *tiles, const TileGen &gen, int64_t ntiles, int bmin,
                         hipStream_t s, bool lookback) {
    if (ntiles <= 0) return hipSuccess;
    if (lookback && (tiles || ntiles > UNPACK_LB_MAX_CHUNKS)) return hipErrorInvalidValue;
    const size_t lds = unpack_lds_bytes(bmin < UNPACK_MIN_B ? UNPACK_MIN_B : bmin);
    if (lookback) hipLaunchKernelGGL(k_unpack<true>, dim3((unsigned)ntiles), dim3(UNPACK_THREADS), lds, s, jobs, tiles, gen);
    else hipLaunchKernelGGL(k_unpack<false>, dim3((unsigned)ntiles), dim3(UNPACK_THREADS), lds, s, jobs, tiles, gen);
    return hipGetLastError();
}

hipError_t launch_carry_fix(const DevJob *jobs, const Tile *tiles, const TileGen &gen, int64_t ntiles,
                            hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_carry_fix, dim3((unsigned)((ntiles + CARRY_FIX_WAVES - 1) / CARRY_FIX_WAVES)),
                       dim3(64 * CARRY_FIX_WAVES), 0, s, jobs, tiles, gen, ntiles);
    return hipGetLastError();
}

hipError_t launch_twiddle_stage(int prime, double *out, int m, int64_t Q, int64_t NoverL, const double *tlo,
                                const double *thi, int h, hipStream_t s) {
    const int64_t total = (int64_t)(m - 1) * Q;
    if (total <= 0) return hipSuccess;
    if (prime == 0)
        hipLaunchKernelGGL(k_twiddle_stage<0>, dim3(grid_for(total, 256)), dim3(256), 0, s, out, m, Q, NoverL, tlo,
                           thi, h);
    else
        hipLaunchKernelGGL(k_twiddle_stage<1>, dim3(grid_for(total, 256)), dim3(256), 0, s, out, m, Q, NoverL, tlo,
                           thi, h);
    return hipGetLastError();
}

hipError_t launch_pointwise(int prime, double *xa, const double *xb, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (prime == 0)
        hipLaunchKernelGGL(k_pointwise<0>, dim3(grid_for(n, 256)), dim3(256), 0, s, xa, xb, n);
    else
        hipLaunchKernelGGL(k_pointwise<1>, dim3(grid_for(n, 256)), dim3(256), 0, s, xa, xb, n);
    return hipGetLastError();
}

hipError_t launch_pack(double *x, int64_t N, const uint64_t *limbs, int64_t n, int b, int64_t nch, hipStream_t s) {
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack, dim3(grid_for(N, 256)), dim3(256), 0, s, x, N, limbs, n, b, nch);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_combine(const CombineJob *jobs, int64_t count, hipStream_t s, bool lens) {
    if (count <= 0) return hipSuccess;
    if (lens) {
        hipLaunchKernelGGL(k_combine_init, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, jobs, count);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_combine, dim3((unsigned)count), dim3(COMBINE_THREADS), 0, s, jobs);
    return hipGetLastError();
}

hipError_t launch_basecase(const BaseJob *jobs, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_basecase, dim3((unsigned)count), dim3(UNPACK_THREADS), 0, s, jobs);
    return hipGetLastError();
}

hipError_t launch_basecase_one(const BaseJob &j, hipStream_t s) {
    hipLaunchKernelGGL(k_basecase_one, dim3(1), dim3(UNPACK_THREADS), 0, s, j);
    return hipGetLastError();
}

hipError_t launch_barrett_fix(const BarrettFixJob *jobs, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_barrett_fix, dim3((unsigned)count), dim3(UNPACK_THREADS), 0, s, jobs);
    return hipGetLastError();
}

hipError_t launch_expand_tiles(const TileRange *ranges, int64_t nranges, Tile *tiles, hipStream_t s) {
    if (nranges <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_expand_tiles, dim3((unsigned)nranges), dim3(256), 0, s, ranges, tiles);
    return hipGetLastError();
}

hipError_t launch_top_limbs(const DevJob *jobs, int64_t count, uint64_t *out, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_top_limbs, dim3(grid_for(count, 256)), dim3(256), 0, s, jobs, count, out);
    return hipGetLastError();
}

hipError_t set_lds_limit(size_t bytes) {
    hipError_t e = hipFuncSetAttribute((const void *)k_group_fwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void *)k_group_inv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void *)k_group_fwd_w6, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void *)k_group_inv_w6, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    // k_small also holds both plans' stage descriptors in static LDS
    e = hipFuncSetAttribute((const void *)k_small, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(bytes - 2 * SMALL_MAX_STAGES * sizeof(StageDesc)));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void *)k_mid, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace nbh
