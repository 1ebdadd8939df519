// e^x for the reaction-rate exponents of the FP64 chemistry (k_f = A T^b exp(-E/(Ru T)) and the
// equilibrium factor exp(sum_s nu_s G_s/RT), src/Reaction.cpp:522-545), callable from host code
// as well, so that its accuracy is checked against the C library on the CPU.
//
// x = n ln2 + r with n = rint(x log2 e) and |r| <= ln2/2; ln2 is split in two (fdlibm's ln2_hi,
// whose low bits are zero so that n ln2_hi is exact for |n| < 2^11, and ln2_lo), which keeps r
// accurate to about 2^-60 relative. e^r comes from its Taylor series to r^13 (the truncation is
// below 2^-56 of the result), evaluated by Horner's rule; ldexp applies 2^n, which also yields
// the infinity or zero (or denormal) of results out of range. A NaN argument gives NaN; infinite
// arguments, which a physical cell never produces, give NaN too. Against the C library's exp over
// [-708, 709] it is within one ulp, equal in about 94 % of arguments.
//
// It replaces the general device exp in the kinetics: that routine's explicit range selects
// (two compares and three selects per call) are what ldexp already covers for finite
// arguments, and the kinetics evaluates 34 exponentials per cell and sub-step.
#pragma once

#if defined(__HIPCC__)
#define EU3D_HD __host__ __device__
#else
#define EU3D_HD
#endif

namespace eu3d {

EU3D_HD inline double exp_fin(double x) {
    constexpr double LOG2E = 1.4426950408889634;
    constexpr double LN2_HI = 6.93147180369123816490e-01;
    constexpr double LN2_LO = 1.90821492927058770002e-10;
    const double n = __builtin_rint(x * LOG2E);
    double r = __builtin_fma(-n, LN2_HI, x);
    r = __builtin_fma(-n, LN2_LO, r);
    double p = 1.0 / 6227020800.0;   // 1/13!
    p = __builtin_fma(p, r, 1.0 / 479001600.0);
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)n);
}

}  // namespace eu3d
