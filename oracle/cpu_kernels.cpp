// Pair-term kernel of the CPU baseline: overlap value and SE(3) force sums of one pose.
//
// TEST INFRASTRUCTURE ONLY (the oracle package): the CPU baseline timed beside the GPU by
// bench.py and checked against the FP64 oracle by tests; never part of the product path.
//
// Restates the structure of the reference's numba SoA kernels (accel/kernels/cpu_soa.py:27-106,
// _overlap_grad_kernel_soa / _overlap_grad_esp_kernel_soa; fastmath, SVML exp): the ref cloud is
// held as structure-of-arrays so the inner loop over ref points is contiguous and vectorises,
// every pair term is FP32 with one vector exp (glibc's libmvec here, as SVML is for numba), the
// per-fit-point partial sums are FP32 and are folded into FP64 totals once per fit point.
// This translation unit alone is compiled with -ffast-math (numba: fastmath on the kernels, not
// on the optimiser tail); the kernel is cloned for AVX-512 and AVX2 and picked at load time.
#include <math.h>
#include <stdint.h>

namespace {

// out[0] = sum E, out[1..3] = sum_b f_b, out[4..12] = sum_b f_b (x) b0, with
// f_b = sum_a E_ab (a - b'), E_ab = exp(-alpha/2 |a - b'|^2 [- (c_a - c_b)^2 / lam]).
template <bool ESP>
__attribute__((always_inline)) inline void pose_sums(const float* __restrict ax, const float* __restrict ay, const float* __restrict az,
               const float* __restrict ac, int n, const float* __restrict fit, int m, const float* R,
               const float* t, float half_alpha, float inv_lam, double* out) {
  double V = 0.0, g[3] = {0.0, 0.0, 0.0}, G[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < m; ++b) {
    const float* p = fit + 4 * b;
    const float bx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + t[0];
    const float by = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + t[1];
    const float bz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + t[2];
    const float cb = p[3];
    float v = 0.f, fx = 0.f, fy = 0.f, fz = 0.f;
#pragma omp simd reduction(+ : v, fx, fy, fz)
    for (int a = 0; a < n; ++a) {
      const float dx = ax[a] - bx, dy = ay[a] - by, dz = az[a] - bz;
      float arg = -half_alpha * (dx * dx + dy * dy + dz * dz);
      if (ESP) {
        const float dc = ac[a] - cb;
        arg -= dc * dc * inv_lam;
      }
      // below -87 the result is under FLT_MIN and adds nothing to an FP32 sum; clamping keeps
      // every lane on libmvec's vector path, which calls the scalar expf for underflowing lanes
      const float e = expf(fmaxf(arg, -87.f));
      v += e;
      fx += e * dx;
      fy += e * dy;
      fz += e * dz;
    }
    V += v;
    g[0] += fx; g[1] += fy; g[2] += fz;
    const double f[3] = {fx, fy, fz};
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) G[3 * i + k] += f[i] * (double)p[k];
  }
  out[0] = V;
  for (int i = 0; i < 3; ++i) out[1 + i] = g[i];
  for (int i = 0; i < 9; ++i) out[4 + i] = G[i];
}

}  // namespace

extern "C" {

__attribute__((target_clones("avx512f", "default")))
void fss_cpu_pose_sums(const float* ax, const float* ay, const float* az, const float* ac, int n,
                       const float* fit, int m, const float* R, const float* t, float half_alpha,
                       float inv_lam, int esp, double* out) {
  if (esp) pose_sums<true>(ax, ay, az, ac, n, fit, m, R, t, half_alpha, inv_lam, out);
  else pose_sums<false>(ax, ay, az, ac, n, fit, m, R, t, half_alpha, inv_lam, out);
}

// The vector width of the clone the loader picks for fss_cpu_pose_sums (reported with the timing).
int fss_cpu_vector_bits(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") ? 512 : 256;
}

}  // extern "C"
