// nbh_field.h -- exact FP64 modular arithmetic for the two NTT primes.
//
// Restates the reference's field layer (src/fpntt.jl:81-121) for host and
// gfx950 device code.  Values are exact integers held in doubles, balanced
// (|x| < a few p).  The results are bit-identical to the reference's because
// every operation is an IEEE double +, -, * or fma in round-to-nearest and the
// sequence of rounding-sensitive operations is the reference's:
//   fp_round(v)   = (v + 1.5*2^52) - 1.5*2^52      exact nearest integer, |v| <= 2^51
//                   (device code: one v_rndne_f64, the same integer -- see fp_round)
//   fp_reduce(x)  = fma(round(x * fl(1/p)), -p, x)
//   fp_mulmod3    = h = x*y; l = fma(x, y, -h); reduce(h) + l
//   fp_mulmod4    = h = x*w; l = fma(x, w, -h); q = round(x*wpinv); fma(q, -p, h) + l
// This file must be compiled with -ffp-contract=off and without fast-math:
// contracting x*y into a neighbouring add would break the fma(x, y, -h)
// error-free product, and reassociation would fold the magic round.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define NBH_HD __host__ __device__ __forceinline__
#else
#define NBH_HD inline
#endif

namespace nbh {

// p1 = 65205*2^34 + 1, p2 = 4095*2^38 + 1 (src/fpntt.jl:78-79)
constexpr uint64_t P1_U = 65205ull * (1ull << 34) + 1;
constexpr uint64_t P2_U = 4095ull * (1ull << 38) + 1;
constexpr double P1_D = 1120213370142721.0;
constexpr double P2_D = 1125625028935681.0;
// fl(1/p), the reference's inv(P); constant-folded with IEEE round-to-nearest
// and checked against 1.0/p at library init.
constexpr double P1_INV = 1.0 / P1_D;
constexpr double P2_INV = 1.0 / P2_D;
constexpr double FP_MAGIC = 6755399441055744.0;  // 1.5 * 2^52
// Garner constant g = p1^-1 mod p2 (208 * p1 = 207 * p2 + 1) and fl(g / p2)
constexpr double GARNER_G = 208.0;
constexpr double GARNER_GP = 208.0 / P2_D;

template <int K> struct Prime;
template <> struct Prime<0> {
    static constexpr double p = P1_D;
    static constexpr double pinv = P1_INV;
    static constexpr uint64_t pu = P1_U;
};
template <> struct Prime<1> {
    static constexpr double p = P2_D;
    static constexpr double pinv = P2_INV;
    static constexpr uint64_t pu = P2_U;
};

// Nearest integer, ties to even (src/fpntt.jl:81-83).  For |v| <= 2^51 the
// magic-constant sum rounds v to the nearest integer with ties to even (1.5*2^52
// is even and its ulp is 1), which is exactly what the rndne instruction
// returns; the device uses the one instruction instead of two adds.  The two
// can differ only in the sign of a zero result (rint(-0.3) is -0.0, the sum
// gives +0.0): the value of every later result is the same, so spectra compare
// equal to the reference's as numbers and every product limb is bit-identical.
NBH_HD double fp_round(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_rint(v);
#else
    return (v + FP_MAGIC) - FP_MAGIC;
#endif
}

template <int K> NBH_HD double fp_reduce(double x) {
    double q = fp_round(x * Prime<K>::pinv);
    return __builtin_fma(q, -Prime<K>::p, x);
}

template <int K> NBH_HD double fp_mulmod3(double x, double y) {
    double h = x * y;
    double l = __builtin_fma(x, y, -h);
    return fp_reduce<K>(h) + l;
}

template <int K> NBH_HD double fp_mulmod4(double x, double w, double wpinv) {
    double h = x * w;
    double l = __builtin_fma(x, w, -h);
    double q = fp_round(x * wpinv);
    return __builtin_fma(q, -Prime<K>::p, h) + l;
}

template <int K> NBH_HD double fp_minbal(double x) {
    constexpr double half = Prime<K>::p / 2;
    return x > half ? x - Prime<K>::p : (x < -half ? x + Prime<K>::p : x);
}

// A prime as a runtime (wave-uniform) value: the kernels take it as an
// argument instead of a template parameter, so both primes run the same
// machine code (half the instruction-cache footprint).  The operations and
// their order are those above, bit for bit.
struct Fld {
    double p, pinv;
};
NBH_HD Fld fld(int k) { return k ? Fld{P2_D, P2_INV} : Fld{P1_D, P1_INV}; }
NBH_HD double fp_reduce(const Fld &f, double x) {
    double q = fp_round(x * f.pinv);
    return __builtin_fma(q, -f.p, x);
}
NBH_HD double fp_mulmod3(const Fld &f, double x, double y) {
    double h = x * y;
    double l = __builtin_fma(x, y, -h);
    return fp_reduce(f, h) + l;
}

// Runtime-prime forms for host plan building.
NBH_HD double fp_reduce_k(double x, int k) { return k ? fp_reduce<1>(x) : fp_reduce<0>(x); }
NBH_HD double fp_mulmod3_k(double x, double y, int k) { return k ? fp_mulmod3<1>(x, y) : fp_mulmod3<0>(x, y); }
NBH_HD double fp_mulmod4_k(double x, double w, double wp, int k) {
    return k ? fp_mulmod4<1>(x, w, wp) : fp_mulmod4<0>(x, w, wp);
}
NBH_HD double fp_minbal_k(double x, int k) { return k ? fp_minbal<1>(x) : fp_minbal<0>(x); }

}  // namespace nbh
