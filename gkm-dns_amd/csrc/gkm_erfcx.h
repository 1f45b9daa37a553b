// gkm_erfcx.h — written by gen_erfcx.py; do not edit.
// erfcx(x) = exp(x^2) erfc(x) on [0, kErfcxB] as a polynomial in u = 2x/kErfcxB - 1
// (Chebyshev interpolant at the first-kind points, converted to monomials in
// 50-digit arithmetic, rounded once), degree 18, truncation error below 2^-24.
#pragma once

namespace gkm {

constexpr double kErfcxB = 6.0;
constexpr int kErfcxDegF = 18;

// coefficients of u^0 .. u^deg
__device__ constexpr float kErfcxF[19] = {
    1.79001151181e-1f,
    -1.6311691919e-1f,
    1.42959419165e-1f,
    -1.2093633144e-1f,
    9.90614640449e-2f,
    -7.89068396307e-2f,
    6.09926471528e-2f,
    -4.47739812847e-2f,
    3.31260044782e-2f,
    -2.90128071185e-2f,
    1.97644707911e-2f,
    -3.09865178875e-4f,
    1.4572574306e-3f,
    -2.15856084251e-2f,
    1.27054871102e-2f,
    1.08308795639e-2f,
    -6.03602528103e-3f,
    -5.80023276809e-3f,
    3.35640187038e-3f,
};

}  // namespace gkm
