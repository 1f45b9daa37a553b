// gkm_erfcx.h — written by gen_erfcx.py; do not edit.
// erfcx(x) = exp(x^2) erfc(x) ~ P(x)/Q(x) on [0, kErfcxCap], deg P = 4, deg Q = 5, Q(0) = 1
// (linearised least squares with Lawson reweighting; relative error of the fit 8.1e-09).
#pragma once

namespace gkm {

constexpr double kErfcxCap = 12.0;
constexpr int kErfcxDegP = 4;
constexpr int kErfcxDegQ = 5;

// coefficients of x^0 .. x^deg
__device__ constexpr float kErfcxP[5] = {1.0f, 1.2527027130126953f, 0.7466817498207092f, 0.23644600808620453f, 0.03432772308588028f};
__device__ constexpr float kErfcxQ[6] = {1.0f, 2.381082773208618f, 2.4334335327148438f, 1.3535279035568237f, 0.4191097617149353f, 0.0608438178896904f};

}  // namespace gkm
