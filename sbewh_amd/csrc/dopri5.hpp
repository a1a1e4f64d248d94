// Dormand-Prince 5(4) tableau (Dormand & Prince 1980; Boost.Odeint's
// runge_kutta_dopri5), for host and device code.
#pragma once
#include <hip/hip_runtime.h>

namespace sbe {

constexpr double A2 = 1.0 / 5, A3 = 3.0 / 10, A4 = 4.0 / 5, A5 = 8.0 / 9;
constexpr double B21 = 1.0 / 5;
constexpr double B31 = 3.0 / 40, B32 = 9.0 / 40;
constexpr double B41 = 44.0 / 45, B42 = -56.0 / 15, B43 = 32.0 / 9;
constexpr double B51 = 19372.0 / 6561, B52 = -25360.0 / 2187, B53 = 64448.0 / 6561, B54 = -212.0 / 729;
constexpr double B61 = 9017.0 / 3168, B62 = -355.0 / 33, B63 = 46732.0 / 5247, B64 = 49.0 / 176,
             B65 = -5103.0 / 18656;
constexpr double C1 = 35.0 / 384, C3 = 500.0 / 1113, C4 = 125.0 / 192, C5 = -2187.0 / 6784, C6 = 11.0 / 84;
constexpr double DC1 = C1 - 5179.0 / 57600, DC3 = C3 - 7571.0 / 16695, DC4 = C4 - 393.0 / 640,
             DC5 = C5 - -92097.0 / 339200, DC6 = C6 - 187.0 / 2100, DC7 = -1.0 / 40;

// time comparisons of integrate_const (Boost.Odeint's less_with_sign /
// less_eq_with_sign for a positive step: machine epsilon as the margin)
__host__ __device__ inline bool less_with_sign(double t1, double t2) { return t2 - t1 > 2.220446049250313e-16; }
__host__ __device__ inline bool less_eq_with_sign(double t1, double t2) { return t1 - t2 <= 2.220446049250313e-16; }

}  // namespace sbe
