// erfc and the Ewald Gaussian from one exponential (real-space Ewald pair term).
//
// erfc(x) = exp(-x^2) t P(t),  t = 1 / (1 + p x),  P(t) = c0 + c1 t + ... + c7 t^7.
// The form is that of Abramowitz & Stegun 7.1.26 (whose five coefficients give
// 1.5e-7 absolute but 4.5e-3 relative error at x = 4); the eight coefficients
// and p here are this library's own fit of P(t) to erfcx(x) = exp(x^2) erfc(x)
// on [0, 5] (weighted least squares iterated toward the minimax solution):
// relative error below 1.7e-8 in exact arithmetic, below 4e-7 when evaluated in
// FP32 (tests/test_erfc_fit.py checks both bounds from these constants).  The
// pair kernel needs exp(-a^2 r^2) for the force anyway, so erfc costs one
// reciprocal and eight multiply-adds more.
#pragma once

#include <hip/hip_runtime.h>

namespace mdx {

constexpr float kErfcP = 4.132480023e-01f;
constexpr float kErfcC[8] = {2.327270672e-01f, 2.397415234e-01f, 1.691467518e-01f, 3.381993625e-01f, -2.537105670e-01f, 5.687582878e-01f, -3.729711517e-01f, 7.810874258e-02f};

// erfc(x) for x >= 0 given g = exp(-x^2)
__host__ __device__ inline float erfc_from_gauss(float x, float g) {
    const float t = 1.0f / (1.0f + kErfcP * x);
    float s = kErfcC[7];
#pragma unroll
    for (int k = 6; k >= 0; --k) s = s * t + kErfcC[k];
    return g * t * s;
}

}  // namespace mdx
