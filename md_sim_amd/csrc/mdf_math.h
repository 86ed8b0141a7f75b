// FP64 elementary functions of the real-space pair kernel: e^x for x <= 0, 1/sqrt(x) and 1/x,
// each within about an ulp of the correctly rounded result at a fraction of the cost of the
// general library sequences.  Only inline device code and one constant table: a translation
// unit that includes this header gets its own copy.
#pragma once
#include <hip/hip_runtime.h>

namespace mdf {

// 2^(j/64) for j = 0..63, each the double nearest the exact value (from 60-digit decimals)
static __constant__ double EXP2_64[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};

// copy the 2^(j/64) table into a block's LDS (every thread of the block calls it; a barrier
// must follow before exp_neg reads the copy)
__device__ __forceinline__ void exp2_table_load(double *lds) {
    for (int k = threadIdx.x; k < 64; k += blockDim.x) lds[k] = EXP2_64[k];
}

// e^x for x <= 0 (the erfc and Buckingham arguments of the pair kernel), FP64, by Tang's
// table method (ACM TOMS 15, 1989): x = (64 m + j) ln2/64 + r with n = 64 m + j =
// rint(64 x / ln2), |r| <= ln2/128; ln2/64 is split into a 21-bit head (n times it is exact for
// |n| < 2^32) and a tail (Cody and Waite's reduction); e^r by its Taylor series to degree 5,
// whose truncation error at |r| = ln2/128 is 3.5e-17 relative; e^x = 2^m 2^(j/64) e^r, the
// table entry read from LDS.  Within about 1.5 ulp.  Arguments below -800 give 0 (as e^-800
// does in FP64); NaN passes through.
__device__ __forceinline__ double exp_neg(double x, const double *tab) {
    x = x < -800.0 ? -800.0 : x;
    const double n = rint(x * 0x1.71547652b82fep+6);   // 64 / ln2
    double r = __builtin_fma(-n, 0x1.62e42p-7, x);
    r = __builtin_fma(-n, 0x1.fdf473de6af28p-28, r);
    double p = 1.0 / 120.0;
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    const int ni = (int)n;
    return ldexp(p * tab[ni & 63], ni >> 6);
}

// 1/sqrt(x) and 1/x from the hardware estimates (v_rsq_f64, v_rcp_f64, relative error e0 about
// 2^-22 or less) refined by one third-order step each, whose error is of order e0^3; the last
// fused multiply-add rounds once, so the results are within about an ulp of the correctly
// rounded ones.
//   1/sqrt(x): y0 = x^-1/2 (1 + d), e = 1 - x y0^2; x^-1/2 = y0 (1 - e)^-1/2
//              = y0 (1 + e/2 + 3 e^2/8 + O(e^3));
//   1/x:       y0 = (1 + d)/x, e = 1 - x y0 = -d; 1/x = y0 / (1 - e) = y0 (1 + e + e^2 + O(e^3)).
__device__ __forceinline__ double rsqrt_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = __builtin_fma(-(x * y), y, 1.0);
    const double c = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(y * e, c, y);
}
__device__ __forceinline__ double rcp_nr(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    const double e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, __builtin_fma(e, e, e), y);
}

}  // namespace mdf
