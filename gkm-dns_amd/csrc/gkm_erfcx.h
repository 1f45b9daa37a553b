// gkm_erfcx.h — written by gen_erfcx.py; do not edit.
// erfcx(x) = exp(x^2) erfc(x) ~ P(x)/Q(x), Q(0) = 1, linearised least squares with Lawson reweighting:
//   FP32: [0, kErfcxCap],   deg P = 4, deg Q = 5, relative error of the fit 8.1e-09;
//   FP64: [0, kErfcxCap64], deg P = 9, deg Q = 10, relative error of the fit 5.1e-17.
#pragma once

namespace gkm {

constexpr double kErfcxCap = 12.0;
constexpr int kErfcxDegP = 4;
constexpr int kErfcxDegQ = 5;
constexpr double kErfcxCap64 = 27.0;
constexpr int kErfcxDegP64 = 9;
constexpr int kErfcxDegQ64 = 10;

// coefficients of x^0 .. x^deg
__device__ constexpr float kErfcxP[5] = {1.0f, 1.2527027130126953f, 0.7466817498207092f, 0.23644600808620453f, 0.03432772308588028f};
__device__ constexpr float kErfcxQ[6] = {1.0f, 2.381082773208618f, 2.4334335327148438f, 1.3535279035568237f, 0.4191097617149353f, 0.0608438178896904f};
__device__ constexpr double kErfcxP64[10] = {1.0, 2.1905849407044062, 2.3739598964475443, 1.6344713618144149, 0.7801764125327235, 0.26664631853053156, 0.065246140487141, 0.011059621834218858, 0.0011873694442432825, 6.241141625050529e-05};
__device__ constexpr double kErfcxQ64[11] = {1.0, 3.318964107799905, 5.119009852027332, 4.843944105251382, 3.123672145935521, 1.4395972264485621, 0.4823643216003878, 0.1166980516035428, 0.019657979990497263, 0.0021045575438454864, 0.00011062135507402754};

}  // namespace gkm
