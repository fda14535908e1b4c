/*
 * fpntt_oracle.c -- CPU restatement of NativeBigInt.jl's FP64 two-prime NTT
 * multiplication engine (src/fpntt.jl), used ONLY as test infrastructure:
 * the parity checker for the HIP library and the CPU baseline timed beside it.
 * Nothing in the product library links, loads or calls this file.
 *
 * Every function cites the reference file:line it follows.  The arithmetic is
 * the reference's exactly: the same primes, generators, Winograd constants,
 * minimal-balanced twiddle tables, stage order and per-butterfly
 * reduce/mulmod schedule, so the spectra it produces are the reference's
 * doubles.  Build with -ffp-contract=off and without fast-math: exactness
 * rests on the FMA error-free product and the magic-constant round.
 *
 * Pinning (see tests/test_oracle.py): field ops against Python big-int
 * arithmetic over the reference's lazy-range edge sets; twiddles and
 * Winograd constants against closed forms; the transform against a brute
 * force DFT, the round-trip identity and O(N^2) cyclic convolution; ntt_len
 * against the reference test's literal values and brute force; unpack against
 * big-int CRT; products against libgmp and Python ints.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <pthread.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------
 * Per-prime context (src/fpntt.jl:50-79)
 * p1 = 65205*2^34+1 (p-1 = 2^34 3^4 5 7 23), p2 = 4095*2^38+1 (2^38 3^2 5 7 13)
 */
typedef struct {
    uint64_t pi;       /* prime as an integer */
    double p;          /* prime as a double (exact) */
    double pinv;       /* fl(1/p): inv(P) in the reference */
    int two_adic;      /* 2-adicity of p-1 */
    uint64_t gen;      /* smallest generator in 3..63 */
    int nfac;
    uint64_t facs[6];  /* distinct prime factors of p-1 */
} ctx_t;

static ctx_t CTX[2];
static pthread_once_t ctx_once = PTHREAD_ONCE_INIT;

static uint64_t powmod_u(uint64_t b, uint64_t e, uint64_t m) {
    u128 r = 1, x = b % m;
    while (e) {
        if (e & 1) r = (r * x) % m;
        x = (x * x) % m;
        e >>= 1;
    }
    return (uint64_t)r;
}

/* modular inverse by extended Euclid (Base.invmod in the reference) */
static uint64_t invmod_u(uint64_t a, uint64_t m) {
    __int128 t = 0, nt = 1, r = m, nr = a % m;
    while (nr) {
        __int128 q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += m;
    return (uint64_t)t;
}

/* fp_generator (src/fpntt.jl:56-62) */
static uint64_t fp_generator(uint64_t p, const uint64_t *facs, int nfac) {
    for (uint64_t g = 3; g <= 63; g++) {
        int ok = 1;
        for (int i = 0; i < nfac; i++)
            if (powmod_u(g, (p - 1) / facs[i], p) == 1) { ok = 0; break; }
        if (ok) return g;
    }
    return 0;
}

static void ctx_init(void) {
    static const uint64_t f1[] = {2, 3, 5, 7, 23};
    static const uint64_t f2[] = {2, 3, 5, 7, 13};
    CTX[0].pi = 65205ULL * (1ULL << 34) + 1;
    CTX[0].two_adic = 34;
    CTX[0].nfac = 5;
    memcpy(CTX[0].facs, f1, sizeof f1);
    CTX[1].pi = 4095ULL * (1ULL << 38) + 1;
    CTX[1].two_adic = 38;
    CTX[1].nfac = 5;
    memcpy(CTX[1].facs, f2, sizeof f2);
    for (int i = 0; i < 2; i++) {
        CTX[i].p = (double)CTX[i].pi;
        CTX[i].pinv = 1.0 / CTX[i].p;
        CTX[i].gen = fp_generator(CTX[i].pi, CTX[i].facs, CTX[i].nfac);
    }
}

static const ctx_t *ctx_get(int k) {
    pthread_once(&ctx_once, ctx_init);
    return &CTX[k];
}

/* ------------------------------------------------------------------------
 * Field arithmetic (src/fpntt.jl:81-121)
 */
#define FP_MAGIC 6755399441055744.0 /* 1.5 * 2^52 */

/* fp_round (src/fpntt.jl:83): exact round-to-nearest for |v| <= 2^51 */
static inline double fp_round(double v) { return (v + FP_MAGIC) - FP_MAGIC; }

/* fp_reduce (src/fpntt.jl:88-91) */
static inline double fp_reduce(double x, const ctx_t *F) {
    double q = fp_round(x * F->pinv);
    return fma(q, -F->p, x);
}

/* fp_mulmod 3-arg (src/fpntt.jl:101-105) */
static inline double fp_mulmod3(double x, double y, const ctx_t *F) {
    double h = x * y;
    double l = fma(x, y, -h);
    return fp_reduce(h, F) + l;
}

/* fp_mulmod 4-arg, Shoup quotient (src/fpntt.jl:106-111) */
static inline double fp_mulmod4(double x, double w, double wpinv, const ctx_t *F) {
    double h = x * w;
    double l = fma(x, w, -h);
    double q = fp_round(x * wpinv);
    return fma(q, -F->p, h) + l;
}

/* fp_minbal (src/fpntt.jl:118-121) */
static inline double fp_minbal(double x, const ctx_t *F) {
    double half = F->p / 2;
    return x > half ? x - F->p : (x < -half ? x + F->p : x);
}

/* ------------------------------------------------------------------------
 * Winograd constants (src/fpntt.jl:127-168)
 */
static void fp_dft5_consts(uint64_t c, const ctx_t *F, double *k) {
    double inv2 = (double)invmod_u(2, F->pi);
    double c1 = fp_reduce((double)c, F);
    double c2 = fp_mulmod3(c1, c1, F), c3 = fp_mulmod3(c2, c1, F), c4 = fp_mulmod3(c3, c1, F);
    double inv4 = fp_mulmod3(inv2, inv2, F);
    k[0] = fp_minbal(-inv4, F);
    k[1] = fp_minbal(fp_mulmod3((c1 + c4) - (c2 + c3), inv4, F), F);
    k[2] = fp_minbal(fp_mulmod3(c1 - c4, inv2, F), F);
    k[3] = fp_minbal(fp_mulmod3(c2 - c3, inv2, F), F);
}

static void fp_dft7_consts(uint64_t c, const ctx_t *F, double *k) {
    uint64_t p = F->pi;
    double inv2 = (double)invmod_u(2, p);
    double inv3 = (double)invmod_u(3, p);
    double cj[7];
    for (int j = 1; j <= 6; j++) cj[j] = (double)powmod_u(c, (uint64_t)j, p);
#define A_(j) fp_mulmod3(cj[j] + cj[7 - (j)], inv2, F)
#define B_(j) fp_mulmod3(cj[j] - cj[7 - (j)], inv2, F)
    double a0 = A_(1), a1 = A_(3), a2 = A_(2);
    double b0 = B_(1), b1 = B_(3), b2 = B_(2);
#undef A_
#undef B_
    double bp0 = fp_minbal(fp_mulmod3(a0 - a2, inv3, F), F);
    double bp1 = fp_minbal(fp_mulmod3(a1 - a2, inv3, F), F);
    double n0 = fp_minbal(fp_mulmod3(b0 - b2, inv3, F), F);
    double n1 = fp_minbal(fp_mulmod3(b1 + b2, inv3, F), F);
    k[0] = fp_minbal(fp_mulmod3((a0 + a1) + a2, inv3, F), F);
    k[1] = bp0;
    k[2] = bp1;
    k[3] = fp_minbal(bp0 + bp1, F);
    k[4] = fp_minbal(fp_mulmod3((b0 + b2) - b1, inv3, F), F);
    k[5] = n0;
    k[6] = n1;
    k[7] = fp_minbal(n0 + n1, F);
}

/* fp_twiddles (src/fpntt.jl:173-186): minimal balanced r^0..r^(n-1),
 * generated as an 8-lane chain stepped by the Shoup mulmod with r^8. */
static double *fp_twiddles(uint64_t r, int64_t n, const ctx_t *F) {
    int64_t np = (n + 7) / 8 * 8;
    double *t = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 8));
    double r8 = fp_reduce((double)powmod_u(r, 8, F->pi), F);
    double r8p = r8 / F->p;
    double v[8];
    for (int k = 0; k < 8; k++) v[k] = fp_reduce((double)powmod_u(r, (uint64_t)k, F->pi), F);
    for (int64_t j = 0; j < np; j += 8) {
        for (int k = 0; k < 8; k++) {
            v[k] = fp_minbal(v[k], F);
            t[j + k] = v[k];
            v[k] = fp_mulmod4(v[k], r8, r8p, F);
        }
    }
    return t;
}

/* ------------------------------------------------------------------------
 * Plans (src/fpntt.jl:191-271)
 */
typedef struct {
    int m;          /* radix */
    int64_t span;   /* block length the stage runs over */
    int64_t Q;      /* span / m */
    double rot[8];  /* combine constants (radix-4: rot[0] = wi) */
    double *tw[6];  /* m-1 twiddle rows of Q entries each */
} stage_t;

typedef struct {
    int prime;      /* 0 or 1 */
    int64_t N, N2;
    double ninv, ninvp;
    int nst;
    stage_t st[40];
} plan_t;

/* build_fp_4 (src/fpntt.jl:199-211), without the <8-lane padding the vector
 * small-q path needs: a scalar butterfly indexes j < q directly. */
static void build_fp_4(stage_t *s, const double *table, int64_t N2, int64_t L, double wi) {
    int64_t q = L >> 2, st = N2 / L;
    s->m = 4; s->span = N2; s->Q = q; s->rot[0] = wi;
    for (int t = 0; t < 3; t++) s->tw[t] = (double *)malloc(sizeof(double) * (size_t)q);
    for (int64_t j = 0; j < q; j++) {
        s->tw[0][j] = table[j * st];
        s->tw[1][j] = table[2 * j * st];
        s->tw[2][j] = table[3 * j * st];
    }
}

/* build_fp_odd (src/fpntt.jl:213-225) */
static void build_fp_odd(stage_t *s, int m, int64_t span, uint64_t root, const ctx_t *F) {
    int64_t Q = span / m;
    uint64_t c = powmod_u(root, (uint64_t)Q, F->pi);
    s->m = m; s->span = span; s->Q = Q;
    if (m == 5) fp_dft5_consts(c, F, s->rot);
    else if (m == 7) fp_dft7_consts(c, F, s->rot);
    else
        for (int r = 1; r <= m - 1; r++)
            s->rot[r - 1] = fp_reduce((double)powmod_u(c, (uint64_t)r, F->pi), F);
    for (int r = 1; r <= m - 1; r++) {
        uint64_t wr = powmod_u(root, (uint64_t)r, F->pi);
        s->tw[r - 1] = fp_twiddles(wr, Q, F);
    }
}

static int tz64(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }

/* build_fp_plan (src/fpntt.jl:236-271).  Returns NULL for an invalid N. */
plan_t *or_plan_build(int64_t N, int prime) {
    const ctx_t *F = ctx_get(prime);
    if (N < 1) return NULL;
    int k = tz64((uint64_t)N);
    int64_t m = N >> k;
    if (!(m == 1 || m == 3 || m == 5 || m == 7 || m == 9 || m == 15 || m == 21 || m == 35 ||
          m == 45 || m == 63 || m == 105 || m == 315))
        return NULL;
    if (!(m == 1 || k >= 2) || k > F->two_adic) return NULL;
    if ((F->pi - 1) % (uint64_t)N != 0) return NULL;
    uint64_t w = powmod_u(F->gen, (F->pi - 1) / (uint64_t)N, F->pi);

    plan_t *P = (plan_t *)calloc(1, sizeof(plan_t));
    P->prime = prime;
    P->N = N;
    int64_t span = N;
    uint64_t r = w;
    static const int mfs[4] = {7, 5, 3, 3};
    for (int i = 0; i < 4; i++) {
        int mf = mfs[i];
        if (m % mf == 0) {
            build_fp_odd(&P->st[P->nst++], mf, span, r, F);
            r = powmod_u(r, (uint64_t)mf, F->pi);
            span /= mf;
            m /= mf;
        }
    }
    double *tw = fp_twiddles(r, span, F);
    double wi = span >= 4 ? fp_reduce((double)powmod_u(r, (uint64_t)(span >> 2), F->pi), F) : 1.0;
    for (int64_t L = span; L >= 4; L >>= 2) build_fp_4(&P->st[P->nst++], tw, span, L, wi);
    free(tw);
    P->N2 = span;
    P->ninv = fp_reduce((double)invmod_u((uint64_t)N % F->pi, F->pi), F);
    P->ninvp = P->ninv / F->p;
    return P;
}

void or_plan_free(plan_t *P) {
    if (!P) return;
    for (int i = 0; i < P->nst; i++)
        for (int t = 0; t < P->st[i].m - 1; t++) free(P->st[i].tw[t]);
    free(P);
}

/* ------------------------------------------------------------------------
 * Stages (src/fpntt.jl:305-701).  Scalar lane form of each butterfly; the
 * reference's 8-lane and small-q shuffle paths run the same per-butterfly
 * arithmetic on the same (j, twiddle) pairs.
 */

/* fp_dft3 (src/fpntt.jl:310-313) */
#define DFT3(a, b, c, w3)                            \
    do {                                             \
        double u_ = fp_mulmod3((b) - (c), (w3), F);  \
        double y0_ = (a) + ((b) + (c));              \
        double y1_ = ((a) - (c)) + u_;               \
        double y2_ = ((a) - (b)) - u_;               \
        a = y0_; b = y1_; c = y2_;                   \
    } while (0)

/* The stage loops below run over the butterfly index j with one pointer per
 * leg: the legs are disjoint runs of Q points, so each loop vectorizes over j
 * (8 doubles a vector in the AVX-512 build, as the reference's VF8 kernels do)
 * with the same per-lane IEEE operations as the scalar loop. */

/* fp_radix3! (src/fpntt.jl:315-354) */
static void radix3(double *x, int64_t o, const stage_t *s, const ctx_t *F, int fwd) {
    const int64_t Q = s->Q;
    const double w3 = s->rot[0];
    const ctx_t Fl = *F;
    const double *restrict t1 = s->tw[0], *restrict t2 = s->tw[1];
    double *restrict x0 = x + o, *restrict x1 = x0 + Q, *restrict x2 = x1 + Q;
#define F (&Fl)
#pragma GCC ivdep
    for (int64_t j = 0; j < Q; j++) {
        double y0 = x0[j], y1 = x1[j], y2 = x2[j];
        if (fwd) DFT3(y0, y1, y2, w3);
        y0 = fp_reduce(y0, F);
        y1 = fp_mulmod3(y1, t1[j], F);
        y2 = fp_mulmod3(y2, t2[j], F);
        if (!fwd) DFT3(y0, y1, y2, w3);
        x0[j] = y0; x1[j] = y1; x2[j] = y2;
    }
#undef F
}

/* fp_dft5 (src/fpntt.jl:371-382) */
static inline void dft5(double *y, const double *k, const ctx_t *F) {
    double x0 = y[0], x1 = y[1], x2 = y[2], x3 = y[3], x4 = y[4];
    double t1 = x1 + x4, t3 = x1 - x4;
    double t2 = x2 + x3, t4 = x2 - x3;
    double S = t1 + t2, D = t1 - t2;
    double y0 = x0 + S;
    double r = fp_reduce(x0 + fp_mulmod3(S, k[0], F), F);
    double q = fp_mulmod3(D, k[1], F);
    double g1 = fp_mulmod3(t3, k[2], F) + fp_mulmod3(t4, k[3], F);
    double g2 = fp_mulmod3(t3, k[3], F) - fp_mulmod3(t4, k[2], F);
    double u = r + q, v = r - q;
    y[0] = y0; y[1] = u + g1; y[2] = v + g2; y[3] = v - g2; y[4] = u - g1;
}

/* fp_radix5! (src/fpntt.jl:384-442) */
static void radix5(double *x, int64_t o, const stage_t *s, const ctx_t *F, int fwd) {
    const int64_t Q = s->Q;
    const ctx_t Fl = *F;
    double rot[8];
    memcpy(rot, s->rot, sizeof rot);
    const double *restrict t1 = s->tw[0], *restrict t2 = s->tw[1], *restrict t3 = s->tw[2], *restrict t4 = s->tw[3];
    double *restrict x0 = x + o, *restrict x1 = x0 + Q, *restrict x2 = x1 + Q, *restrict x3 = x2 + Q,
           *restrict x4 = x3 + Q;
#define F (&Fl)
#pragma GCC ivdep
    for (int64_t j = 0; j < Q; j++) {
        double y[5] = {x0[j], x1[j], x2[j], x3[j], x4[j]};
        if (fwd) dft5(y, rot, F);
        y[0] = fp_reduce(y[0], F);
        y[1] = fp_mulmod3(y[1], t1[j], F);
        y[2] = fp_mulmod3(y[2], t2[j], F);
        y[3] = fp_mulmod3(y[3], t3[j], F);
        y[4] = fp_mulmod3(y[4], t4[j], F);
        if (!fwd) {
            dft5(y, rot, F);
            y[0] = fp_reduce(y[0], F);
        }
        x0[j] = y[0]; x1[j] = y[1]; x2[j] = y[2]; x3[j] = y[3]; x4[j] = y[4];
    }
#undef F
}

/* fp_dft7 (src/fpntt.jl:459-492) */
static inline void dft7(double *y, const double *rot, const ctx_t *F, int red) {
    double x0 = y[0];
    double s1 = y[1] + y[6], d1 = y[1] - y[6];
    double s2 = y[2] + y[5], d2 = y[2] - y[5];
    double s3 = y[3] + y[4], d3 = y[3] - y[4];
    if (red) {
        s1 = fp_reduce(s1, F); s2 = fp_reduce(s2, F); s3 = fp_reduce(s3, F);
        d1 = fp_reduce(d1, F); d2 = fp_reduce(d2, F); d3 = fp_reduce(d3, F);
    }
    double S = (s1 + s2) + s3;
    double y0 = x0 + S;
    double m1 = fp_mulmod3(S, rot[0], F);
    double va = s1 - s3, vb = s2 - s3;
    double w0 = fp_mulmod3(va, rot[1], F);
    double w1 = fp_mulmod3(vb, rot[2], F);
    double w2 = fp_mulmod3(va + vb, rot[3], F);
    double mn = fp_mulmod3((d1 + d2) - d3, rot[4], F);
    double wa = d1 + d3, wb = -(d2 + d3);
    double q0 = fp_mulmod3(wa, rot[5], F);
    double q1 = fp_mulmod3(wb, rot[6], F);
    double q2 = fp_mulmod3(wa + wb, rot[7], F);
    double u1 = fp_reduce(m1 + (3 * w0 - w2), F);
    double u3 = fp_reduce((m1 + 2 * w2) - 3 * (w0 + w1), F);
    double u2 = fp_reduce(m1 + (3 * w1 - w2), F);
    double r0 = q0 - q1, r1 = q2 - q0;
    double v1 = fp_reduce(mn + (2 * r0 + r1), F);
    double v3 = fp_reduce((r0 + 2 * r1) - mn, F);
    double v2 = fp_reduce((mn - r0) + r1, F);
    double a1 = x0 + u1, a2 = x0 + u2, a3 = x0 + u3;
    y[0] = y0; y[1] = a1 + v1; y[2] = a2 + v2; y[3] = a3 + v3;
    y[4] = a3 - v3; y[5] = a2 - v2; y[6] = a1 - v1;
}

/* fp_radix7! (src/fpntt.jl:494-537) */
static void radix7(double *x, int64_t o, const stage_t *s, const ctx_t *F, int fwd) {
    const int64_t Q = s->Q;
    const ctx_t Fl = *F;
    double rot[8];
    memcpy(rot, s->rot, sizeof rot);
    const double *restrict t1 = s->tw[0], *restrict t2 = s->tw[1], *restrict t3 = s->tw[2], *restrict t4 = s->tw[3],
                 *restrict t5 = s->tw[4], *restrict t6 = s->tw[5];
    double *restrict x0 = x + o, *restrict x1 = x0 + Q, *restrict x2 = x1 + Q, *restrict x3 = x2 + Q,
           *restrict x4 = x3 + Q, *restrict x5 = x4 + Q, *restrict x6 = x5 + Q;
#define F (&Fl)
#pragma GCC ivdep
    for (int64_t j = 0; j < Q; j++) {
        double y[7] = {x0[j], x1[j], x2[j], x3[j], x4[j], x5[j], x6[j]};
        if (fwd) dft7(y, rot, F, 0);
        y[0] = fp_reduce(y[0], F);
        y[1] = fp_mulmod3(y[1], t1[j], F);
        y[2] = fp_mulmod3(y[2], t2[j], F);
        y[3] = fp_mulmod3(y[3], t3[j], F);
        y[4] = fp_mulmod3(y[4], t4[j], F);
        y[5] = fp_mulmod3(y[5], t5[j], F);
        y[6] = fp_mulmod3(y[6], t6[j], F);
        if (!fwd) {
            dft7(y, rot, F, 1);
            y[0] = fp_reduce(y[0], F);
        }
        x0[j] = y[0]; x1[j] = y[1]; x2[j] = y[2]; x3[j] = y[3]; x4[j] = y[4]; x5[j] = y[5]; x6[j] = y[6];
    }
#undef F
}

/* fp_dft4 (src/fpntt.jl:543-549) */
#define DFT4(a, b, c, d, wi)                              \
    do {                                                  \
        double apc_ = (a) + (c), amc_ = (a) - (c);        \
        double bpd_ = (b) + (d);                          \
        double ibmd_ = fp_mulmod3((b) - (d), (wi), F);    \
        a = apc_ + bpd_; b = amc_ + ibmd_;                \
        c = apc_ - bpd_; d = amc_ - ibmd_;                \
    } while (0)

/* fp_radix4! (src/fpntt.jl:645-701), all three Q regimes */
static void radix4(double *x, int64_t o, const stage_t *s, const ctx_t *F, int fwd) {
    const int64_t Q = s->Q;
    const double wi = s->rot[0];
    const ctx_t Fl = *F;
    const double *restrict t1 = s->tw[0], *restrict t2 = s->tw[1], *restrict t3 = s->tw[2];
#define F (&Fl)
    for (int64_t b0 = o; b0 < o + s->span; b0 += 4 * Q) {
        double *restrict x0 = x + b0, *restrict x1 = x0 + Q, *restrict x2 = x1 + Q, *restrict x3 = x2 + Q;
#pragma GCC ivdep
        for (int64_t j = 0; j < Q; j++) {
            double y0 = x0[j], y1 = x1[j], y2 = x2[j], y3 = x3[j];
            if (fwd) DFT4(y0, y1, y2, y3, wi);
            y0 = fp_reduce(y0, F);
            y1 = fp_mulmod3(y1, t1[j], F);
            y2 = fp_mulmod3(y2, t2[j], F);
            y3 = fp_mulmod3(y3, t3[j], F);
            if (!fwd) DFT4(y0, y1, y2, y3, wi);
            x0[j] = y0; x1[j] = y1; x2[j] = y2; x3[j] = y3;
        }
    }
#undef F
}

/* fp_radix2! (src/fpntt.jl:630-638) */
static void radix2(double *x, int64_t N, const ctx_t *F) {
    for (int64_t s = 0; s < N; s += 2) {
        double u = x[s], v = x[s + 1];
        x[s] = fp_reduce(u + v, F);
        x[s + 1] = fp_reduce(u - v, F);
    }
}

/* fp_stage! (src/fpntt.jl:703-713) */
static void stage_run(double *x, const stage_t *s, const plan_t *P, int fwd) {
    const ctx_t *F = ctx_get(P->prime);
    if (s->m == 4) {           /* radix-4 stages span the whole N2 block */
        for (int64_t o = 0; o < P->N; o += s->span) radix4(x, o, s, F, fwd);
        return;
    }
    for (int64_t o = 0; o < P->N; o += s->span) {
        if (s->m == 3) radix3(x, o, s, F, fwd);
        else if (s->m == 5) radix5(x, o, s, F, fwd);
        else radix7(x, o, s, F, fwd);
    }
}

/* fp_ntt_fwd! (src/fpntt.jl:715-721) */
void or_fwd(double *x, const plan_t *P) {
    for (int i = 0; i < P->nst; i++) stage_run(x, &P->st[i], P, 1);
    if (tz64((uint64_t)P->N2) & 1) radix2(x, P->N, ctx_get(P->prime));
}

/* fp_ntt_rev! (src/fpntt.jl:727-733) */
void or_rev(double *x, const plan_t *P) {
    if (tz64((uint64_t)P->N2) & 1) radix2(x, P->N, ctx_get(P->prime));
    for (int i = P->nst - 1; i >= 0; i--) stage_run(x, &P->st[i], P, 0);
}

/* ------------------------------------------------------------------------
 * Coefficient layer (src/fpntt.jl:744-785)
 */

/* fp_ntt_pack! (src/fpntt.jl:744-768): chunk i = bits [i*b, i*b+b) of the
 * n-limb operand (bits past limb n-1 read as zero), zero fill to N. */
void or_pack(double *x, int64_t N, const uint64_t *limbs, int64_t n, int b, int64_t nch) {
    uint64_t mask = (b >= 64) ? ~0ULL : ((1ULL << b) - 1);
    for (int64_t i = 0; i < nch; i++) {
        int64_t bit = i * b;
        int64_t w = bit >> 6;
        int sh = (int)(bit & 63);
        uint64_t c = limbs[w] >> sh;
        if (sh && w + 1 < n) c |= limbs[w + 1] << (64 - sh);
        x[i] = (double)(c & mask);
    }
    for (int64_t i = nch; i < N; i++) x[i] = 0.0;
}

/* fp_ntt_pointwise! (src/fpntt.jl:770-785) */
void or_pointwise(double *xa, const double *xb, int64_t n, int prime) {
    const ctx_t *F = ctx_get(prime);
    for (int64_t i = 0; i < n; i++) xa[i] = fp_mulmod3(xa[i], xb[i], F);
}

/* addmul_1! semantics (src/kernels/mul.jl:32-70): r += a*b, returns carry */
uint64_t or_addmul_1(uint64_t *r, const uint64_t *a, int64_t n, uint64_t b) {
    uint64_t c = 0;
    for (int64_t i = 0; i < n; i++) {
        u128 p = (u128)a[i] * b + r[i] + c;
        r[i] = (uint64_t)p;
        c = (uint64_t)(p >> 64);
    }
    return c;
}

/* fp_unpack_coeff (src/fpntt.jl:806-816) */
static inline void unpack_coeff(const double *x1, const double *x2, int64_t idx,
                                double n1, double n1p, double n2, double n2p,
                                double g, double gp, uint64_t *c1o, uint64_t *uuo) {
    const ctx_t *F1 = ctx_get(0), *F2 = ctx_get(1);
    double v1 = fp_mulmod4(x1[idx], n1, n1p, F1);
    v1 = v1 < 0 ? v1 + F1->p : v1;
    double v2 = fp_mulmod4(x2[idx], n2, n2p, F2);
    double u = fp_mulmod4(v2 - fp_reduce(v1, F2), g, gp, F2);
    *c1o = (uint64_t)v1;
    *uuo = (uint64_t)(u < 0 ? u + F2->p : u);
}

/* fp_ntt_unpack2! (src/fpntt.jl:869-925): Garner CRT of coefficient i read at
 * index (N-i) mod N, scaled by 1/N, streamed as S1 = sum c1_i 2^(b i) into r
 * and S2 = sum u_i 2^(b i) through two 128-bit windows, then r += p1*S2. */
void or_unpack2(uint64_t *r, int64_t rn, const double *x1, const double *x2, int64_t N,
                int64_t nconv, int b, double n1, double n1p, double n2, double n2p) {
    const ctx_t *F1 = ctx_get(0), *F2 = ctx_get(1);
    double g = (double)invmod_u(F1->pi % F2->pi, F2->pi);
    double gp = g / F2->p;
    uint64_t *s2 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(rn > 0 ? rn : 1));
    u128 w1 = 0, w2 = 0;
    int64_t outw = 0;
    int s = 0;
    for (int64_t c = 0; c < nconv; c++) {
        int64_t idx = c == 0 ? 0 : N - c;
        uint64_t c1, uu;
        unpack_coeff(x1, x2, idx, n1, n1p, n2, n2p, g, gp, &c1, &uu);
        w1 += (u128)c1 << s;
        w2 += (u128)uu << s;
        s += b;
        if (s >= 64) {
            if (outw < rn) { r[outw] = (uint64_t)w1; s2[outw] = (uint64_t)w2; }
            w1 >>= 64; w2 >>= 64;
            outw++;
            s -= 64;
        }
    }
    for (; outw < rn; outw++) {
        r[outw] = (uint64_t)w1; s2[outw] = (uint64_t)w2;
        w1 >>= 64; w2 >>= 64;
    }
    or_addmul_1(r, s2, rn, F1->pi);
    free(s2);
}

/* ------------------------------------------------------------------------
 * Sizing (src/fpntt.jl:933-977, src/kernels/addsub.jl:150)
 */
static int top_set_bit64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }
static int top_set_bit128(u128 x) {
    uint64_t hi = (uint64_t)(x >> 64);
    return hi ? 64 + top_set_bit64(hi) : top_set_bit64((uint64_t)x);
}

/* magnitude_bits (src/kernels/addsub.jl:150) */
int64_t or_magnitude_bits(const uint64_t *a, int64_t n) {
    return 64 * (n - 1) + top_set_bit64(a[n - 1]);
}

/* ntt_len (src/fpntt.jl:933-951); returns -1 where the reference throws */
int64_t or_ntt_len(int64_t T) {
    static const int64_t ms_small[] = {1, 3, 5, 9};
    static const int64_t ms_mid[] = {1, 3, 5, 7, 9, 15, 21, 105};
    static const int64_t ms_large[] = {1, 3, 5, 7, 9, 15, 21, 35, 45, 63, 105, 315};
    const int64_t *ms;
    int nms;
    if (T <= (1LL << 14)) { ms = ms_small; nms = 4; }
    else if (T <= (1LL << 20)) { ms = ms_mid; nms = 8; }
    else { ms = ms_large; nms = 12; }
    int maxk = ctx_get(0)->two_adic < ctx_get(1)->two_adic ? ctx_get(0)->two_adic : ctx_get(1)->two_adic;
    int64_t best = INT64_MAX;
    for (int i = 0; i < nms; i++) {
        int64_t m = ms[i];
        int64_t cl = (T + m - 1) / m;
        int k = top_set_bit64((uint64_t)(cl - 1));
        if (k < 2) k = 2;
        if (k > maxk) continue;
        int64_t c = m << k;
        if (c < best) best = c;
    }
    return best == INT64_MAX ? -1 : best;
}

/* fp_ntt_params (src/fpntt.jl:964-977) */
int or_ntt_params(int64_t bits_a, int64_t bits_b, int *b_out, int64_t *N_out) {
    const ctx_t *F1 = ctx_get(0), *F2 = ctx_get(1);
    u128 pp = (u128)F1->pi * F2->pi - 1;
    uint64_t pmin = F1->pi < F2->pi ? F1->pi : F2->pi;
    int maxb = top_set_bit64(pmin) - 1;
    int64_t mn = bits_a < bits_b ? bits_a : bits_b;
    int nclow = top_set_bit64((uint64_t)((mn + maxb - 1) / maxb));
    int bstart = (top_set_bit128(pp) - nclow + 3) >> 1;
    if (bstart > maxb) bstart = maxb;
    for (int b = bstart; b >= 1; b--) {
        int64_t nca = (bits_a + b - 1) / b, ncb = (bits_b + b - 1) / b;
        int64_t mnc = nca < ncb ? nca : ncb;
        u128 d = (u128)((1ULL << b) - 1) * ((1ULL << b) - 1);
        if ((u128)mnc <= pp / d) {
            *b_out = b;
            *N_out = or_ntt_len(nca + ncb - 1);
            return *N_out < 0 ? -1 : 0;
        }
    }
    return -2;
}

/* ------------------------------------------------------------------------
 * Engine entry points (src/fpntt.jl:982-1037).  Plans are cached per (N, prime)
 * like fp_ntt_plan (src/fpntt.jl:273-293): read lock-free, insert under a lock.
 */
#define PLAN_CACHE_MAX 256
static plan_t *plan_cache[PLAN_CACHE_MAX];
static int plan_cache_n;
static pthread_mutex_t plan_lock = PTHREAD_MUTEX_INITIALIZER;

static const plan_t *plan_get(int64_t N, int prime) {
    int n = __atomic_load_n(&plan_cache_n, __ATOMIC_ACQUIRE);
    for (int i = 0; i < n; i++)
        if (plan_cache[i]->N == N && plan_cache[i]->prime == prime) return plan_cache[i];
    pthread_mutex_lock(&plan_lock);
    n = plan_cache_n;
    for (int i = 0; i < n; i++)
        if (plan_cache[i]->N == N && plan_cache[i]->prime == prime) {
            pthread_mutex_unlock(&plan_lock);
            return plan_cache[i];
        }
    plan_t *P = or_plan_build(N, prime);
    if (P && n < PLAN_CACHE_MAX) {
        plan_cache[n] = P;
        __atomic_store_n(&plan_cache_n, n + 1, __ATOMIC_RELEASE);
    }
    pthread_mutex_unlock(&plan_lock);
    return P;
}

const plan_t *or_plan(int64_t N, int prime) { return plan_get(N, prime); }

/* mul_fpntt2! (src/fpntt.jl:982-1014): r[0..m+n) = a[0..m) * b[0..n).
 * Returns 0, or -1 for an unsupported length (the reference's ArgumentError). */
int or_mul(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n) {
    int64_t bits_a = or_magnitude_bits(a, m), bits_b = or_magnitude_bits(b, n);
    int bch;
    int64_t N;
    if (or_ntt_params(bits_a, bits_b, &bch, &N)) return -1;
    const plan_t *P1 = plan_get(N, 0), *P2 = plan_get(N, 1);
    if (!P1 || !P2) return -1;
    int64_t nca = (bits_a + bch - 1) / bch, ncb = (bits_b + bch - 1) / bch;
    double *buf = (double *)malloc(sizeof(double) * (size_t)(3 * N));
    double *t1 = buf, *t2 = buf + N, *t3 = buf + 2 * N;
    or_pack(t3, N, a, m, bch, nca);
    memcpy(t1, t3, sizeof(double) * (size_t)N);
    or_fwd(t1, P1);
    or_pack(t2, N, b, n, bch, ncb);
    or_fwd(t2, P1);
    or_pointwise(t1, t2, N, 0);
    or_rev(t1, P1);
    or_fwd(t3, P2);
    or_pack(t2, N, b, n, bch, ncb);
    or_fwd(t2, P2);
    or_pointwise(t3, t2, N, 1);
    or_rev(t3, P2);
    or_unpack2(r, m + n, t1, t3, N, nca + ncb - 1, bch, P1->ninv, P1->ninvp, P2->ninv, P2->ninvp);
    free(buf);
    return 0;
}

/* sqr_fpntt2! (src/fpntt.jl:1016-1037) */
int or_sqr(uint64_t *r, const uint64_t *a, int64_t n) {
    int64_t bits = or_magnitude_bits(a, n);
    int bch;
    int64_t N;
    if (or_ntt_params(bits, bits, &bch, &N)) return -1;
    const plan_t *P1 = plan_get(N, 0), *P2 = plan_get(N, 1);
    if (!P1 || !P2) return -1;
    int64_t nca = (bits + bch - 1) / bch;
    double *buf = (double *)malloc(sizeof(double) * (size_t)(2 * N));
    double *t1 = buf, *t2 = buf + N;
    or_pack(t2, N, a, n, bch, nca);
    memcpy(t1, t2, sizeof(double) * (size_t)N);
    or_fwd(t1, P1);
    or_pointwise(t1, t1, N, 0);
    or_rev(t1, P1);
    or_fwd(t2, P2);
    or_pointwise(t2, t2, N, 1);
    or_rev(t2, P2);
    or_unpack2(r, 2 * n, t1, t2, N, 2 * nca - 1, bch, P1->ninv, P1->ninvp, P2->ninv, P2->ninvp);
    free(buf);
    return 0;
}

/* ------------------------------------------------------------------------
 * Flat accessors for the ctypes wrapper (oracle/oracle.py)
 */
uint64_t or_prime(int k) { return ctx_get(k)->pi; }
uint64_t or_gen(int k) { return ctx_get(k)->gen; }
double or_pinv(int k) { return ctx_get(k)->pinv; }

void or_vec_reduce(double *out, const double *x, int64_t n, int k) {
    const ctx_t *F = ctx_get(k);
    for (int64_t i = 0; i < n; i++) out[i] = fp_reduce(x[i], F);
}
void or_vec_mulmod3(double *out, const double *x, const double *y, int64_t n, int k) {
    const ctx_t *F = ctx_get(k);
    for (int64_t i = 0; i < n; i++) out[i] = fp_mulmod3(x[i], y[i], F);
}
void or_vec_mulmod4(double *out, const double *x, const double *w, const double *wp, int64_t n, int k) {
    const ctx_t *F = ctx_get(k);
    for (int64_t i = 0; i < n; i++) out[i] = fp_mulmod4(x[i], w[i], wp[i], F);
}
void or_vec_minbal(double *out, const double *x, int64_t n, int k) {
    const ctx_t *F = ctx_get(k);
    for (int64_t i = 0; i < n; i++) out[i] = fp_minbal(x[i], F);
}
void or_twiddles(double *out, uint64_t r, int64_t n, int k) {
    double *t = fp_twiddles(r, n, ctx_get(k));
    memcpy(out, t, sizeof(double) * (size_t)n);
    free(t);
}

/* plan introspection: counts, then per-stage fields */
int or_plan_nstages(const plan_t *P) { return P->nst; }
int64_t or_plan_N2(const plan_t *P) { return P->N2; }
double or_plan_ninv(const plan_t *P) { return P->ninv; }
double or_plan_ninvp(const plan_t *P) { return P->ninvp; }
void or_plan_stage(const plan_t *P, int i, int *m, int64_t *span, int64_t *Q, double *rot8) {
    const stage_t *s = &P->st[i];
    *m = s->m; *span = s->span; *Q = s->Q;
    memcpy(rot8, s->rot, sizeof(double) * 8);
}
void or_plan_twiddle_row(const plan_t *P, int i, int t, double *out) {
    const stage_t *s = &P->st[i];
    memcpy(out, s->tw[t], sizeof(double) * (size_t)s->Q);
}
