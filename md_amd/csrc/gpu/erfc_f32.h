// Generated by fit_erfc.py (this directory); do not edit by hand.
//
// erfc(x) = exp(-x^2) t Q(t), t = 1 / (1 + p x), x >= 0: the Abramowitz &
// Stegun 7.1.26 form with a degree-7 Q fitted for minimax RELATIVE error of
// t Q(t) against erfc(x) exp(x^2) on [0, 5.0] (Lawson's algorithm).
// Fit error 1.50e-08.  Largest relative error by x range of t Q(t) in FP32
// against erfc(x) exp(x^2), and of the whole formula in FP32 against FP64
// erfc (the second adds the rounding of x^2 in exp(-x^2), unbiased):
//   [0.0, 1.0]  4.94e-07  5.64e-07
//   [1.0, 2.0]  3.95e-07  4.47e-07
//   [2.0, 3.0]  3.08e-07  7.64e-07
//   [3.0, 3.5]  3.09e-07  7.93e-07
//   [3.5, 4.0]  3.13e-07  8.12e-07
//   [4.0, 5.0]  3.08e-07  1.28e-06
#ifndef MD_AMD_GPU_ERFC_F32_H
#define MD_AMD_GPU_ERFC_F32_H

namespace mdg {

constexpr float kErfcP = 3.199999928e-01F;

// erfc(x) for 0 <= x <= 5.0 from e = exp(-x^2)
__device__ __forceinline__ float erfc_f32(float x, float e) {
    const float t = __builtin_amdgcn_rcpf(fmaf(kErfcP, x, 1.0f));
    float q = -1.767021567e-01F;
    q = fmaf(q, t, 7.018226385e-01F);
    q = fmaf(q, t, -7.590268254e-01F);
    q = fmaf(q, t, 7.832862735e-01F);
    q = fmaf(q, t, -1.580047160e-01F);
    q = fmaf(q, t, 2.611013949e-01F);
    q = fmaf(q, t, 1.659615636e-01F);
    q = fmaf(q, t, 1.815618575e-01F);
    return q * t * e;
}

}  // namespace mdg

#endif
