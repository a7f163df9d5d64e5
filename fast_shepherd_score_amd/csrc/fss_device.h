// Device-side building blocks shared by every gfx950 kernel of libfss_hip.so.
//
// Math contract (restated from score/analytical_gradients/_torch.py):
//   overlap  O = K * sum_a sum_b w_ab * exp(-alpha/2 |a - b'|^2),  b' = R b + t,  K = (pi/(2 alpha))^1.5
//   dO/dt    = K alpha sum w E (a - b')                         (:495-501)
//   dO/dR    = K alpha sum w E (a - b') b^T                     (:503-512)
//   ESP      w_ab = exp(-(c_a - c_b)^2 / lam)                   (:544-566)
//   pharm    same-type pairs only, alpha/K/category from the type; directional weight
//            w = (clamp(D,0,1)+2)/3 (cat 1) or (|D|+2)/3 (cat 2), D = (R v_fit).v_ref, with the
//            weight gradient (1/3) K E c v_ref v_fit^T          (:223-361)
//            extended points (cat 4, 5): anchor weight 1 or 0, plus K exp(-alpha/2 |e_a - e_b'|^2)
//            over the extended points e = anchor + unit vector  (:288-357)
//   rotation R(q_hat) in (w,x,y,z) order and its four Jacobians (:28-139)
//
// Shape channel arithmetic runs on pre-scaled coordinates: with s = sqrt(alpha/2 * log2(e)) every
// coordinate (and the translation) is multiplied by s while it is staged in LDS and the ESP value
// by sqrt(log2(e)/lam), so the Gaussian and the ESP weight are one exp2 of an expanded quadratic
// form (see shape_accum).  The sums are rescaled once per pose: dO/dt = K alpha / s * sum(f~),
// dO/dR = K alpha / s^2 * G~.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/fss.h"

namespace fss {

constexpr float kLog2e = 1.4426950408889634f;
constexpr double kLog2eD = 1.4426950408889634;

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// Rotation from a raw quaternion in FP64: qn = |q|, qh = q / qn, R(qh) row-major.
__device__ __forceinline__ void rotation_fp64(const float q[4], double qh[4], double& qn,
                                              double Rd[9], double* inv_qn = nullptr) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  qn = sqrt(w * w + x * x + y * y + z * z);
  const double inv = 1.0 / qn;
  if (inv_qn) *inv_qn = inv;
  qh[0] = w * inv; qh[1] = x * inv; qh[2] = y * inv; qh[3] = z * inv;
  const double a = qh[0], b = qh[1], c = qh[2], d = qh[3];
  Rd[0] = 1.0 - 2.0 * (c * c + d * d); Rd[1] = 2.0 * (b * c - d * a);       Rd[2] = 2.0 * (b * d + c * a);
  Rd[3] = 2.0 * (b * c + d * a);       Rd[4] = 1.0 - 2.0 * (b * b + d * d); Rd[5] = 2.0 * (c * d - b * a);
  Rd[6] = 2.0 * (b * d - c * a);       Rd[7] = 2.0 * (c * d + b * a);       Rd[8] = 1.0 - 2.0 * (b * b + c * c);
}

// g_k = sum_ij G_ij (dR/dq_k)_ij at a unit quaternion qh (ag/_torch.py:62-173).
template <typename T>
__device__ __forceinline__ void project_to_quat(const T G[9], const double qh[4], double g[4]) {
  const double w = qh[0], x = qh[1], y = qh[2], z = qh[3];
  const double G0 = G[0], G1 = G[1], G2 = G[2], G3 = G[3], G4 = G[4], G5 = G[5], G6 = G[6],
               G7 = G[7], G8 = G[8];
  // dR/dw = [[0,-2z,2y],[2z,0,-2x],[-2y,2x,0]]
  g[0] = -2 * z * G1 + 2 * y * G2 + 2 * z * G3 - 2 * x * G5 - 2 * y * G6 + 2 * x * G7;
  // dR/dx = [[0,2y,2z],[2y,-4x,-2w],[2z,2w,-4x]]
  g[1] = 2 * y * G1 + 2 * z * G2 + 2 * y * G3 - 4 * x * G4 - 2 * w * G5 + 2 * z * G6 + 2 * w * G7 - 4 * x * G8;
  // dR/dy = [[-4y,2x,2w],[2x,0,2z],[-2w,2z,-4y]]
  g[2] = -4 * y * G0 + 2 * x * G1 + 2 * w * G2 + 2 * x * G3 + 2 * z * G5 - 2 * w * G6 + 2 * z * G7 - 4 * y * G8;
  // dR/dz = [[-4z,-2w,2x],[2w,-4z,2y],[2x,2y,0]]
  g[3] = -4 * z * G0 - 2 * w * G1 + 2 * x * G2 + 2 * w * G3 - 4 * z * G4 + 2 * y * G5 + 2 * x * G6 + 2 * y * G7;
}

// Per-lane partial sums of one channel for one pose.
struct Acc {
  float V;
  float gt[3];
  float G[9];
  __device__ __forceinline__ void zero() {
    V = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) gt[i] = 0.f;
#pragma unroll
    for (int i = 0; i < 9; ++i) G[i] = 0.f;
  }
};

// Butterfly sum over `width` consecutive lanes (width a power of two <= 64); every lane of the
// group ends with the bit-identical total (fp add is commutative).  Level L adds the value of a
// lane in the other half of the aligned 2^(L+1)-lane block; once the halves hold uniform partial
// sums any such partner gives the xor-butterfly's result, so the levels inside a 16-lane row use
// DPP moves (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror), level 4 a ds_swizzle
// xor 16 and level 5 a permute across the two 32-lane halves.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int L>
__device__ __forceinline__ float sum_level(float v) {
  if constexpr (L == 0) return v + dpp_move<0xB1>(v);
  else if constexpr (L == 1) return v + dpp_move<0x4E>(v);
  else if constexpr (L == 2) return v + dpp_move<0x141>(v);
  else if constexpr (L == 3) return v + dpp_move<0x140>(v);
  else if constexpr (L == 4) return v + __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
  else return v + __shfl_xor(v, 32, 64);
}
__device__ __forceinline__ float group_sum(float v, int width) {
  if (width > 1) v = sum_level<0>(v);
  if (width > 2) v = sum_level<1>(v);
  if (width > 4) v = sum_level<2>(v);
  if (width > 8) v = sum_level<3>(v);
  if (width > 16) v = sum_level<4>(v);
  if (width > 32) v = sum_level<5>(v);
  return v;
}
__device__ __forceinline__ double group_sum(double v, int width) {
  for (int m = 1; m < width; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}
// The 13 sums of a pose level by level, so each level's exchanges are independent.
template <int L>
__device__ __forceinline__ void sum_level(Acc& a, bool grad) {
  a.V = sum_level<L>(a.V);
  if (grad) {
#pragma unroll
    for (int i = 0; i < 3; ++i) a.gt[i] = sum_level<L>(a.gt[i]);
#pragma unroll
    for (int i = 0; i < 9; ++i) a.G[i] = sum_level<L>(a.G[i]);
  }
}
__device__ __forceinline__ void group_sum(Acc& a, int width, bool grad) {
  if (width > 1) sum_level<0>(a, grad);
  if (width > 2) sum_level<1>(a, grad);
  if (width > 4) sum_level<2>(a, grad);
  if (width > 8) sum_level<3>(a, grad);
  if (width > 16) sum_level<4>(a, grad);
  if (width > 32) sum_level<5>(a, grad);
}

// Sum over a group of `width` consecutive lanes for any width (groups need not be aligned to a
// power of two; a group may not straddle the wavefront).  j = the lane's index in its group.
// Fixed tree: at stride s, lanes with j % 2s == 0 add the partial of lane j + s (if in the
// group); the group's first lane then holds the total, which every lane takes back, so all lanes
// of a group end with the same bits whatever the other groups hold.
__device__ __forceinline__ float group_sum_any(float v, int width, int j) {
  const int lane = threadIdx.x & 63;
  for (int s = 1; s < width; s <<= 1) {
    const float o = __shfl(v, lane + s < 64 ? lane + s : lane, 64);
    if ((j & (2 * s - 1)) == 0 && j + s < width) v += o;
  }
  return __shfl(v, lane - j, 64);
}
__device__ __forceinline__ void group_sum_any(Acc& a, int width, int j, bool grad) {
  a.V = group_sum_any(a.V, width, j);
  if (grad) {
#pragma unroll
    for (int i = 0; i < 3; ++i) a.gt[i] = group_sum_any(a.gt[i], width, j);
#pragma unroll
    for (int i = 0; i < 9; ++i) a.G[i] = group_sum_any(a.G[i], width, j);
  }
}
// The same sums through LDS, for a workgroup that is one wavefront of `nl` grouped lanes (lane
// groups of any width): every lane stores its 13 partials (red[13 * lane + k], an odd stride, so
// the stores of a row of lanes hit distinct banks); lane j of a group then adds partial k = j,
// j + width, ... over the group's lanes in lane order and stores the total at
// red[13 * (nl + grp) + k], and every lane of the group loads the 13 totals.  One lane forms each total, so a group's lanes end with the
// same bits.  A wavefront's LDS operations complete in issue order; the wavefront-scope fences
// keep the compiler from moving the loads of other lanes' values across the stores.
__device__ __forceinline__ void group_sum_lds(Acc& a, int width, int j, int grp, int nl, float* red,
                                              bool grad) {
  const int lane = threadIdx.x, n = grad ? 13 : 1;
  float v[13];
  v[0] = a.V;
#pragma unroll
  for (int i = 0; i < 3; ++i) v[1 + i] = a.gt[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) v[4 + i] = a.G[i];
  for (int k = 0; k < n; ++k) red[13 * lane + k] = v[k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const float* part = red + 13 * (lane - j);
  float* tot = red + 13 * (nl + grp);
  for (int k = j; k < n; k += width) {
    float s = part[k];
    for (int i = 1; i < width; ++i) s += part[13 * i + k];
    tot[k] = s;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  a.V = tot[0];
  if (grad) {
#pragma unroll
    for (int i = 0; i < 3; ++i) a.gt[i] = tot[1 + i];
#pragma unroll
    for (int i = 0; i < 9; ++i) a.G[i] = tot[4 + i];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Floats of LDS group_sum_lds takes for `groups` lane groups of `width` lanes: 13 partials per
// lane, 13 totals per group.
constexpr int red_floats(int groups, int width) { return 13 * (groups * width + groups); }

// A pose's sums over a lane group of `width` > 64 lanes, whole wavefronts (the streamed path):
// each wavefront's sums by DPP levels, then every lane adds its group's wavefront totals from LDS
// in wavefront order, so a group's lanes end with the same bits.  Every thread of the workgroup
// calls it (two barriers); red holds 13 floats per wavefront of the workgroup.
__device__ __forceinline__ void pose_sum_waves(Acc& a, int width, float* red) {
  group_sum(a, 64, true);
  const int w = threadIdx.x >> 6, nw = width >> 6, w0 = (threadIdx.x / width) * nw;
  __syncthreads();   // the previous sums' readers are done with red
  if ((threadIdx.x & 63) == 0) {
    float* r = red + 13 * w;
    r[0] = a.V;
#pragma unroll
    for (int i = 0; i < 3; ++i) r[1 + i] = a.gt[i];
#pragma unroll
    for (int i = 0; i < 9; ++i) r[4 + i] = a.G[i];
  }
  __syncthreads();
  Acc s;
  s.zero();
  for (int k = 0; k < nw; ++k) {
    const float* r = red + 13 * (w0 + k);
    s.V += r[0];
#pragma unroll
    for (int i = 0; i < 3; ++i) s.gt[i] += r[1 + i];
#pragma unroll
    for (int i = 0; i < 9; ++i) s.G[i] += r[4 + i];
  }
  a = s;
}

// A pose's sums over its lane group: DPP levels for power-of-two widths; else the LDS exchange
// when the kernel reserved `red` (a one-wavefront workgroup of nl grouped lanes), the shuffle
// tree otherwise.
__device__ __forceinline__ void pose_sum(Acc& a, int width, int j, bool grad, int grp = 0, int nl = 0,
                                         float* red = nullptr) {
  if ((width & (width - 1)) == 0) group_sum(a, width, grad);
  else if (red) group_sum_lds(a, width, j, grp, nl, red, grad);
  else group_sum_any(a, width, j, grad);
}

// ---------------------------------------------------------------------------------------------
// Shape channel, FP32, expanded form on scaled coordinates staged in LDS.
//
// With s = sqrt(alpha/2 log2 e), sc = sqrt(log2 e / lam) and the ref centroid c as origin:
//   a~ = s (a - c),  b~' = R (s b) + s t - s c,  c~ = sc * esp
//   E_ab = exp2(-|a~ - b~'|^2 - (c~_a - c~_b)^2) = exp2(g_ab - D) * exp2(D - h_b)
//   g_ab = W_a + 2 a~ . b~' + 2 c~_a c~_b,   W_a = -|a~|^2 - c~_a^2,   h_b = |b~'|^2 + c~_b^2
// LDS rows: ref4[a] = (2 a~, W_a - D), refc[a] = 2 c~_a (ESP only), fit4[b] = (s b, c~_b).
// A pair term is then 3-4 FMA + v_exp_f32 + 1 add + 3 FMA; exp2(D - h_b) is applied once per fit
// row.  The shift D = kExpShift moves the exponent window: exp2(g_ab - D) = E_ab 2^(h_b - D) and
// |2 a~| E_ab <= 2 (sqrt(h_b) + 1) < 2^5, so a sum over up to 2^14 ref rows (more than LDS holds),
// gradient sums included, stays finite while h_b <= kHMax = D + 108 (about 18 A from the ref
// centroid at alpha 0.81), and a term flushes to zero only where E_ab < 2^-40.  Fit rows beyond
// that radius take the direct form.
// Lane j of a group of `lpp` lanes owns fit rows b = j, j+lpp, ...; two rows per pass over the
// ref cloud so one LDS broadcast read feeds two pair terms.  The group runs ceil(M/lpp) row slots:
// floor of half that in two-row passes (a lane whose second row lies past M runs a copy of its
// first, zero-weighted), plus one single-row pass when the count is odd.
// Output (per lane, scaled units): V = sum E, gt = sum_b f_b, G = sum_b f_b (s b)^T with
// f_b = sum_a E_ab (a~ - b~'_b).
// ---------------------------------------------------------------------------------------------
constexpr float kExpShift = 86.f;
constexpr float kHMax = kExpShift + 108.f;

template <bool ESP>
__device__ __forceinline__ void direct_pair(const float4* __restrict__ ref, const float* __restrict__ refc,
                                            int N, float x, float y, float z, float cb, float& v,
                                            float& fx, float& fy, float& fz) {
  for (int a = 0; a < N; ++a) {
    const float4 A = ref[a];
    const float dx = 0.5f * A.x - x, dy = 0.5f * A.y - y, dz = 0.5f * A.z - z;
    float s2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    if (ESP) {
      const float dc = 0.5f * refc[a] - cb;
      s2 = fmaf(dc, dc, s2);
    }
    const float e = fast_exp2(-s2);
    v += e;
    fx = fmaf(e, dx, fx); fy = fmaf(e, dy, fy); fz = fmaf(e, dz, fz);
  }
}

template <bool GRAD>
__device__ __forceinline__ void accumulate_rows(Acc& acc, const float4& p0, float v0, float fx0, float fy0,
                                                float fz0, const float4& p1, float v1, float fx1, float fy1,
                                                float fz1) {
  acc.V += v0 + v1;
  if (GRAD) {
    acc.gt[0] += fx0 + fx1; acc.gt[1] += fy0 + fy1; acc.gt[2] += fz0 + fz1;
    acc.G[0] = fmaf(fx0, p0.x, fmaf(fx1, p1.x, acc.G[0]));
    acc.G[1] = fmaf(fx0, p0.y, fmaf(fx1, p1.y, acc.G[1]));
    acc.G[2] = fmaf(fx0, p0.z, fmaf(fx1, p1.z, acc.G[2]));
    acc.G[3] = fmaf(fy0, p0.x, fmaf(fy1, p1.x, acc.G[3]));
    acc.G[4] = fmaf(fy0, p0.y, fmaf(fy1, p1.y, acc.G[4]));
    acc.G[5] = fmaf(fy0, p0.z, fmaf(fy1, p1.z, acc.G[5]));
    acc.G[6] = fmaf(fz0, p0.x, fmaf(fz1, p1.x, acc.G[6]));
    acc.G[7] = fmaf(fz0, p0.y, fmaf(fz1, p1.y, acc.G[7]));
    acc.G[8] = fmaf(fz0, p0.z, fmaf(fz1, p1.z, acc.G[8]));
  }
}

// Direct-form twin of shape_accum (the other inner-loop variant): LDS rows ref4[a] = (a~, c~_a),
// fit4[b] = (s b, c~_b); a pair term is 3 sub + 3 FMA/mul (+ 1 sub + 1 FMA for ESP) + v_exp_f32
// (negated operand) + 1 add + 3 FMA.  No radius limit.
template <bool ESP, bool GRAD>
__device__ __forceinline__ void direct_rows(const float4* __restrict__ ref, int N, const float4& p0,
                                            const float4& p1, const float R[9], const float ts[3],
                                            bool two, bool keep1, Acc& acc) {
  const float x0 = fmaf(R[0], p0.x, fmaf(R[1], p0.y, fmaf(R[2], p0.z, ts[0])));
  const float y0 = fmaf(R[3], p0.x, fmaf(R[4], p0.y, fmaf(R[5], p0.z, ts[1])));
  const float z0 = fmaf(R[6], p0.x, fmaf(R[7], p0.y, fmaf(R[8], p0.z, ts[2])));
  float v0 = 0.f, v1 = 0.f;
  float fx0 = 0.f, fy0 = 0.f, fz0 = 0.f, fx1 = 0.f, fy1 = 0.f, fz1 = 0.f;
  if (two) {
    const float x1 = fmaf(R[0], p1.x, fmaf(R[1], p1.y, fmaf(R[2], p1.z, ts[0])));
    const float y1 = fmaf(R[3], p1.x, fmaf(R[4], p1.y, fmaf(R[5], p1.z, ts[1])));
    const float z1 = fmaf(R[6], p1.x, fmaf(R[7], p1.y, fmaf(R[8], p1.z, ts[2])));
#pragma unroll 4
    for (int a = 0; a < N; ++a) {
      const float4 A = ref[a];
      const float dx0 = A.x - x0, dy0 = A.y - y0, dz0 = A.z - z0;
      const float dx1 = A.x - x1, dy1 = A.y - y1, dz1 = A.z - z1;
      float s0 = fmaf(dx0, dx0, fmaf(dy0, dy0, dz0 * dz0));
      float s1 = fmaf(dx1, dx1, fmaf(dy1, dy1, dz1 * dz1));
      if (ESP) {
        const float c0 = A.w - p0.w, c1 = A.w - p1.w;
        s0 = fmaf(c0, c0, s0);
        s1 = fmaf(c1, c1, s1);
      }
      const float e0 = fast_exp2(-s0);
      const float e1 = fast_exp2(-s1);
      v0 += e0;
      v1 += e1;
      if (GRAD) {
        fx0 = fmaf(e0, dx0, fx0); fy0 = fmaf(e0, dy0, fy0); fz0 = fmaf(e0, dz0, fz0);
        fx1 = fmaf(e1, dx1, fx1); fy1 = fmaf(e1, dy1, fy1); fz1 = fmaf(e1, dz1, fz1);
      }
    }
  } else {
#pragma unroll 4
    for (int a = 0; a < N; ++a) {
      const float4 A = ref[a];
      const float dx0 = A.x - x0, dy0 = A.y - y0, dz0 = A.z - z0;
      float s0 = fmaf(dx0, dx0, fmaf(dy0, dy0, dz0 * dz0));
      if (ESP) {
        const float c0 = A.w - p0.w;
        s0 = fmaf(c0, c0, s0);
      }
      const float e0 = fast_exp2(-s0);
      v0 += e0;
      if (GRAD) { fx0 = fmaf(e0, dx0, fx0); fy0 = fmaf(e0, dy0, fy0); fz0 = fmaf(e0, dz0, fz0); }
    }
  }
  if (!keep1) v1 = fx1 = fy1 = fz1 = 0.f;
  accumulate_rows<GRAD>(acc, p0, v0, fx0, fy0, fz0, p1, v1, fx1, fy1, fz1);
}

template <bool ESP, bool GRAD>
__device__ __forceinline__ void shape_accum_direct(const float4* __restrict__ ref, int N,
                                                   const float4* __restrict__ fit, int M, int j, int lpp,
                                                   const float R[9], const float ts[3], Acc& acc) {
  const int cmax = (M + lpp - 1) / lpp, passes = cmax >> 1;
  int b0 = j;
  for (int k = 0; k < passes; ++k, b0 += 2 * lpp) {
    const bool ok1 = b0 + lpp < M;
    direct_rows<ESP, GRAD>(ref, N, fit[b0], fit[ok1 ? b0 + lpp : b0], R, ts, true, ok1, acc);
  }
  if ((cmax & 1) && b0 < M) {
    const float4 p0 = fit[b0];
    direct_rows<ESP, GRAD>(ref, N, p0, p0, R, ts, false, false, acc);
  }
}

// One fit row against the whole ref cloud, expanded form (the tail pass of shape_accum).
template <bool ESP, bool GRAD>
__device__ __forceinline__ void expanded_row(const float4* __restrict__ ref, const float* __restrict__ refc,
                                             int N, float x, float y, float z, float c, float& v,
                                             float& fx, float& fy, float& fz) {
  const float h = fmaf(x, x, fmaf(y, y, fmaf(z, z, c * c)));
  if (h <= kHMax) {
    float sx = 0.f, sy = 0.f, sz = 0.f, acc = 0.f;
#pragma unroll 4
    for (int a = 0; a < N; ++a) {
      const float4 A = ref[a];
      float g = A.w;
      if (ESP) g = fmaf(refc[a], c, g);
      g = fmaf(A.x, x, fmaf(A.y, y, fmaf(A.z, z, g)));
      const float e = fast_exp2(g);
      acc += e;
      if (GRAD) { sx = fmaf(e, A.x, sx); sy = fmaf(e, A.y, sy); sz = fmaf(e, A.z, sz); }
    }
    const float F = fast_exp2(kExpShift - h);
    v = acc * F;
    if (GRAD) { fx = fmaf(0.5f * F, sx, -x * v); fy = fmaf(0.5f * F, sy, -y * v); fz = fmaf(0.5f * F, sz, -z * v); }
  } else {
    v = fx = fy = fz = 0.f;
    direct_pair<ESP>(ref, refc, N, x, y, z, c, v, fx, fy, fz);
  }
}

template <bool ESP, bool GRAD>
__device__ __forceinline__ void shape_accum(const float4* __restrict__ ref, const float* __restrict__ refc,
                                            int N, const float4* __restrict__ fit, int M, int j, int lpp,
                                            const float R[9], const float ts[3], Acc& acc) {
  const int cmax = (M + lpp - 1) / lpp, passes = cmax >> 1;
  int b0 = j;
  for (int k = 0; k < passes; ++k, b0 += 2 * lpp) {
    const bool ok1 = b0 + lpp < M;
    const float4 p0 = fit[b0];
    const float4 p1 = fit[ok1 ? b0 + lpp : b0];
    const float x0 = fmaf(R[0], p0.x, fmaf(R[1], p0.y, fmaf(R[2], p0.z, ts[0])));
    const float y0 = fmaf(R[3], p0.x, fmaf(R[4], p0.y, fmaf(R[5], p0.z, ts[1])));
    const float z0 = fmaf(R[6], p0.x, fmaf(R[7], p0.y, fmaf(R[8], p0.z, ts[2])));
    const float x1 = fmaf(R[0], p1.x, fmaf(R[1], p1.y, fmaf(R[2], p1.z, ts[0])));
    const float y1 = fmaf(R[3], p1.x, fmaf(R[4], p1.y, fmaf(R[5], p1.z, ts[1])));
    const float z1 = fmaf(R[6], p1.x, fmaf(R[7], p1.y, fmaf(R[8], p1.z, ts[2])));
    const float c0 = ESP ? p0.w : 0.f, c1 = ESP ? p1.w : 0.f;
    const float h0 = fmaf(x0, x0, fmaf(y0, y0, fmaf(z0, z0, c0 * c0)));
    const float h1 = fmaf(x1, x1, fmaf(y1, y1, fmaf(z1, z1, c1 * c1)));
    float v0 = 0.f, v1 = 0.f;
    float fx0 = 0.f, fy0 = 0.f, fz0 = 0.f, fx1 = 0.f, fy1 = 0.f, fz1 = 0.f;
    if (fmaxf(h0, h1) <= kHMax) {
      float sx0 = 0.f, sy0 = 0.f, sz0 = 0.f, sx1 = 0.f, sy1 = 0.f, sz1 = 0.f;
#pragma unroll 4
      for (int a = 0; a < N; ++a) {
        const float4 A = ref[a];
        float g0 = A.w, g1 = A.w;
        if (ESP) {
          const float C = refc[a];
          g0 = fmaf(C, c0, g0);
          g1 = fmaf(C, c1, g1);
        }
        g0 = fmaf(A.x, x0, fmaf(A.y, y0, fmaf(A.z, z0, g0)));
        g1 = fmaf(A.x, x1, fmaf(A.y, y1, fmaf(A.z, z1, g1)));
        const float e0 = fast_exp2(g0);
        const float e1 = fast_exp2(g1);
        v0 += e0;
        v1 += e1;
        if (GRAD) {
          sx0 = fmaf(e0, A.x, sx0); sy0 = fmaf(e0, A.y, sy0); sz0 = fmaf(e0, A.z, sz0);
          sx1 = fmaf(e1, A.x, sx1); sy1 = fmaf(e1, A.y, sy1); sz1 = fmaf(e1, A.z, sz1);
        }
      }
      const float F0 = fast_exp2(kExpShift - h0), F1 = fast_exp2(kExpShift - h1);
      v0 *= F0;
      v1 *= F1;
      if (GRAD) {
        fx0 = fmaf(0.5f * F0, sx0, -x0 * v0); fy0 = fmaf(0.5f * F0, sy0, -y0 * v0); fz0 = fmaf(0.5f * F0, sz0, -z0 * v0);
        fx1 = fmaf(0.5f * F1, sx1, -x1 * v1); fy1 = fmaf(0.5f * F1, sy1, -y1 * v1); fz1 = fmaf(0.5f * F1, sz1, -z1 * v1);
      }
    } else {
      direct_pair<ESP>(ref, refc, N, x0, y0, z0, c0, v0, fx0, fy0, fz0);
      direct_pair<ESP>(ref, refc, N, x1, y1, z1, c1, v1, fx1, fy1, fz1);
    }
    if (!ok1) v1 = fx1 = fy1 = fz1 = 0.f;
    accumulate_rows<GRAD>(acc, p0, v0, fx0, fy0, fz0, p1, v1, fx1, fy1, fz1);
  }
  // single-row pass when the row-slot count is odd
  if ((cmax & 1) && b0 < M) {
    const float4 p0 = fit[b0];
    const float x0 = fmaf(R[0], p0.x, fmaf(R[1], p0.y, fmaf(R[2], p0.z, ts[0])));
    const float y0 = fmaf(R[3], p0.x, fmaf(R[4], p0.y, fmaf(R[5], p0.z, ts[1])));
    const float z0 = fmaf(R[6], p0.x, fmaf(R[7], p0.y, fmaf(R[8], p0.z, ts[2])));
    float v0, fx0, fy0, fz0;
    expanded_row<ESP, GRAD>(ref, refc, N, x0, y0, z0, ESP ? p0.w : 0.f, v0, fx0, fy0, fz0);
    accumulate_rows<GRAD>(acc, p0, v0, fx0, fy0, fz0, p0, 0.f, 0.f, 0.f, 0.f);
  }
}

// ---------------------------------------------------------------------------------------------
// Shape channel on the matrix cores (inner-loop variant 3): the expanded-form exponent
//   D_ab = (W_a - D) + 2 a~ . b~' + 2 c~_a c~_b - m_b
// is a K = 32 product of split-bf16 operands, one v_mfma_f32_16x16x32_bf16 per 16 x 16 tile of
// (ref row, fit row) pairs.  Each FP32 operand x is split exactly into hi + mid + lo bf16 parts
// (x - hi - mid - lo = O(2^-24 x)); for x, y and z the 8 K-slots of a lane group hold the
// products hi.hi, hi.mid, mid.hi, hi.lo, lo.hi, mid.mid, mid.lo, lo.mid (bf16 products are exact
// in FP32, the dropped lo.lo is O(2^-48)); the fourth group holds the ESP products hi.hi, hi.mid,
// mid.hi, the ref constant W_a - D (hi, mid, lo) against 1, and 1 against the column shift
// -m_b (hi, mid).  The result matches an FP32 FMA chain's accuracy.
// The column shift m_b = max(0, h_b - kHMax) keeps every exponent <= kHMax - D for any fit row, so
// this variant has no direct-form fallback; exp2(D + m^_b - h_b) with the same rounded m^_b
// restores the scale once per column.
//
// Operand layouts (v_mfma_f32_16x16x32_bf16, lane l, group g = l / 16, r = l % 16):
//   A: ref row r of the tile, K-slots 8g..8g+7;  B: fit column r, K-slots 8g..8g+7;
//   D: rows 4g..4g+3 of column r.
// LDS (per pair, T = ceil(N / 16) tiles, rows past N have zero coordinates and W = -1e30):
//   amat[64 t + 16 g + r]     8 bf16: A-operand row r of tile t for K-group g
//   xq[12 t + 3 g + k]        float4: 2 a~_k of ref rows 16 t + 4 g + 0..3 (k = x, y, z)
// One wavefront owns one pose; a pass covers 32 fit columns (two B tiles; the second tile's
// MFMA is skipped in a last pass with no more than 16 left) against all T ref tiles, and each
// lane folds the partial sums of one column into the pose sums as the expanded form does per
// fit row.
// ---------------------------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  float r = x - (float)h;
  m = (__bf16)r;
  r -= (float)m;
  l = (__bf16)r;
}

// A-operand K-group row for one ref row: groups 0..2 take 2 a~_k, group 3 the ESP value 2 c~_a and
// the shifted W_a - D.
__device__ __forceinline__ bf16x8 mfma_a_coord(float x) {
  __bf16 h, m, l;
  split3(x, h, m, l);
  bf16x8 a;
  a[0] = h; a[1] = h; a[2] = m; a[3] = h; a[4] = l; a[5] = m; a[6] = m; a[7] = l;
  return a;
}
__device__ __forceinline__ bf16x8 mfma_a_last(float c, float w) {
  __bf16 ch, cm, cl, wh, wm, wl;
  split3(c, ch, cm, cl);
  split3(w, wh, wm, wl);
  const __bf16 one = (__bf16)1.0f;
  bf16x8 a;
  a[0] = ch; a[1] = ch; a[2] = cm; a[3] = wh; a[4] = wm; a[5] = wl; a[6] = one; a[7] = one;
  return a;
}

// The tile loop of one pass: B tile 0 always, B tile 1 when TWO.
template <bool TWO>
__device__ __forceinline__ void mfma_tiles(const float4* __restrict__ am, const float4* __restrict__ xg, int T,
                                           const bf16x8& B0, const bf16x8& B1, float s0[4], float s1[4]) {
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  float v0 = 0.f, sx0 = 0.f, sy0 = 0.f, sz0 = 0.f, v1 = 0.f, sx1 = 0.f, sy1 = 0.f, sz1 = 0.f;
  for (int t = 0; t < T; ++t) {
    const float4 araw = am[64 * t];
    const float4 ax = xg[12 * t], ay = xg[12 * t + 1], az = xg[12 * t + 2];
    const bf16x8 A = __builtin_bit_cast(bf16x8, araw);
    const f32x4 d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B0, zero, 0, 0, 0);
    f32x4 d1 = zero;
    if (TWO) d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B1, zero, 0, 0, 0);
    const float xa[4] = {ax.x, ax.y, ax.z, ax.w}, ya[4] = {ay.x, ay.y, ay.z, ay.w},
                za[4] = {az.x, az.y, az.z, az.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float e0 = fast_exp2(d0[i]);
      v0 += e0;
      sx0 = fmaf(e0, xa[i], sx0); sy0 = fmaf(e0, ya[i], sy0); sz0 = fmaf(e0, za[i], sz0);
      if (TWO) {
        const float e1 = fast_exp2(d1[i]);
        v1 += e1;
        sx1 = fmaf(e1, xa[i], sx1); sy1 = fmaf(e1, ya[i], sy1); sz1 = fmaf(e1, za[i], sz1);
      }
    }
  }
  s0[0] = v0; s0[1] = sx0; s0[2] = sy0; s0[3] = sz0;
  s1[0] = v1; s1[1] = sx1; s1[2] = sy1; s1[3] = sz1;
}

// v_permlane16_swap on a register pair (a, b), 16-lane rows: a <- (a.row0, b.row0, a.row2, b.row2),
// b <- (a.row1, b.row1, a.row3, b.row3).
__device__ __forceinline__ void row_swap(unsigned& a, unsigned& b) {
  const auto s = __builtin_amdgcn_permlane16_swap(a, b, false, false);
  a = s[0];
  b = s[1];
}
__device__ __forceinline__ void row_swap(float& a, float& b) {
  unsigned ua = __float_as_uint(a), ub = __float_as_uint(b);
  row_swap(ua, ub);
  a = __uint_as_float(ua);
  b = __uint_as_float(ub);
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// B-operand K-group of one fit column: a coordinate (groups 0..2) or the ESP value with the
// constant slots (group 3: 1 against W_a - D, the column shift -m_b against 1).
__device__ __forceinline__ u32x4 mfma_b_coord(float x) {
  __bf16 h, m, l;
  split3(x, h, m, l);
  bf16x8 b;
  b[0] = h; b[1] = m; b[2] = h; b[3] = l; b[4] = h; b[5] = m; b[6] = l; b[7] = m;
  return __builtin_bit_cast(u32x4, b);
}
__device__ __forceinline__ u32x4 mfma_b_last(float c, __bf16 mh, __bf16 mm) {
  __bf16 h, m, l;
  split3(c, h, m, l);
  const __bf16 one = (__bf16)1.0f;
  bf16x8 b;
  b[0] = h; b[1] = m; b[2] = h; b[3] = one; b[4] = one; b[5] = one; b[6] = -mh; b[7] = -mm;
  return __builtin_bit_cast(u32x4, b);
}

// One pass covers the 32 fit columns c0 .. c0 + 31 (B tiles 0 and 1).  Lane (g, r) transforms one
// column, c0 + 16 (g & 1) + r, and builds its B operands for the K-groups of its row pair
// {g & 2, (g & 2) + 1}: x and y in rows 0-1, z and the ESP row in rows 2-3.  One permlane16 swap
// per operand dword then hands every lane its own K-group of both tiles (row 2i holds tile 0's
// column, row 2i+1 tile 1's).  After the tile loop the same swap on the (tile 0, tile 1) partial
// sums gives each lane its column's sums over its row pair, so each lane finishes its one column;
// the pose-level butterfly adds the two row pairs.
template <bool ESP>
__device__ __forceinline__ void shape_accum_mfma(const float4* __restrict__ amat, const float4* __restrict__ xq,
                                                 int T, const float4* __restrict__ fit, int M,
                                                 const float R[9], const float ts[3], Acc& acc) {
  const int l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  const bool lo = g < 2;
  const float4* am = amat + 16 * g + r;
  const float4* xg = xq + 3 * g;
  for (int c0 = 0; c0 < M; c0 += 32) {
    // the second tile's MFMA is skipped when no more than 16 columns remain
    const bool two = c0 + 16 < M;
    const int b = c0 + 16 * (g & 1) + r;
    const bool ok = b < M;
    const float4 q = fit[ok ? b : 0];
    const float X = fmaf(R[0], q.x, fmaf(R[1], q.y, fmaf(R[2], q.z, ts[0])));
    const float Y = fmaf(R[3], q.x, fmaf(R[4], q.y, fmaf(R[5], q.z, ts[1])));
    const float Z = fmaf(R[6], q.x, fmaf(R[7], q.y, fmaf(R[8], q.z, ts[2])));
    const float c = ESP ? q.w : 0.f;
    const float h = fmaf(X, X, fmaf(Y, Y, fmaf(Z, Z, c * c)));
    __bf16 mh, mm, ml;
    split3(fmaxf(h - kHMax, 0.f), mh, mm, ml);
    const float F = ok ? fast_exp2(kExpShift + ((float)mh + (float)mm) - h) : 0.f;
    u32x4 B0 = mfma_b_coord(lo ? X : Z);
    u32x4 B1 = lo ? mfma_b_coord(Y) : mfma_b_last(c, mh, mm);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      unsigned a = B0[i], e = B1[i];
      row_swap(a, e);
      B0[i] = a;
      B1[i] = e;
    }
    float s0[4], s1[4];
    const bf16x8 Bt0 = __builtin_bit_cast(bf16x8, B0), Bt1 = __builtin_bit_cast(bf16x8, B1);
    if (two) mfma_tiles<true>(am, xg, T, Bt0, Bt1, s0, s1);
    else mfma_tiles<false>(am, xg, T, Bt0, Bt1, s0, s1);
    // the column finish f_b = F_b (s_b / 2 - b~'_b v_b) is linear in the row sums; a column past
    // M has F = 0, and a one-tile pass leaves zero tile-1 sums
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      row_swap(s0[i], s1[i]);
      s0[i] += s1[i];
    }
    const float w = s0[0] * F, hf = 0.5f * F;
    const float fx = fmaf(hf, s0[1], -X * w), fy = fmaf(hf, s0[2], -Y * w), fz = fmaf(hf, s0[3], -Z * w);
    acc.V += w;
    acc.gt[0] += fx; acc.gt[1] += fy; acc.gt[2] += fz;
    acc.G[0] = fmaf(fx, q.x, acc.G[0]); acc.G[1] = fmaf(fx, q.y, acc.G[1]); acc.G[2] = fmaf(fx, q.z, acc.G[2]);
    acc.G[3] = fmaf(fy, q.x, acc.G[3]); acc.G[4] = fmaf(fy, q.y, acc.G[4]); acc.G[5] = fmaf(fy, q.z, acc.G[5]);
    acc.G[6] = fmaf(fz, q.x, acc.G[6]); acc.G[7] = fmaf(fz, q.y, acc.G[7]); acc.G[8] = fmaf(fz, q.z, acc.G[8]);
  }
}

// ---------------------------------------------------------------------------------------------
// ShaEP surface ESP comparison (FSS_MODE_ESP_COMBO): value and gradient of one pose, FP32.
//   esp = sum over observer points i not inside the other molecule's radius + probe volume of
//         exp(-(pe_i - C sum_m q_m / d_im)^2 / lam)
// for the ref surface against the moved fit atoms (direction 1) and for the moved fit surface
// against the ref atoms (direction 2) (score/electrostatic_scoring.py:128-227).
// Coordinates: ref rows relative to the origin o, fit rows in the fit frame; the pose maps
// x -> R x + tr (tr = t - o).  Direction 1 runs in the fit frame (each observer pulled back by
// R^T), so the fit atoms are read untransformed and its forces are rotated back once.
// Gradient: dE_i/dphi_i = 2 E_i u_i / lam with u_i = pe_i - phi_i; d(1/d)/da = (p - a) / d^3;
// the mask is piecewise constant (zero gradient).  Lanes take observers j, j+lpp, ...; acc gets
// V = esp, gt = dV/dt, G = dV/dR.  rb_* hold (radius + probe)^2.
// ---------------------------------------------------------------------------------------------
struct ShaepK {
  float coul;   // C
  float kq;     // log2(e) / lam
  float c2;     // 2 C / lam
};

template <bool GRAD>
__device__ __forceinline__ void shaep_accum(const float4* __restrict__ srf_r, int Ns,
                                            const float4* __restrict__ atm_f, const float* __restrict__ rb_f,
                                            int Ma, const float4* __restrict__ srf_f, int Ms,
                                            const float4* __restrict__ atm_r, const float* __restrict__ rb_r,
                                            int Na, int j, int lpp, const float R[9], const float tr[3],
                                            const ShaepK k, Acc& acc) {
  float f[3] = {0.f, 0.f, 0.f}, Fm[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int i = j; i < Ns; i += lpp) {
    const float4 P = srf_r[i];
    const float dx = P.x - tr[0], dy = P.y - tr[1], dz = P.z - tr[2];
    const float px = fmaf(R[0], dx, fmaf(R[3], dy, R[6] * dz));
    const float py = fmaf(R[1], dx, fmaf(R[4], dy, R[7] * dz));
    const float pz = fmaf(R[2], dx, fmaf(R[5], dy, R[8] * dz));
    float phi = 0.f;
    bool blocked = false;
    for (int m = 0; m < Ma; ++m) {
      const float4 A = atm_f[m];
      const float ex = px - A.x, ey = py - A.y, ez = pz - A.z;
      const float d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
      blocked |= d2 < rb_f[m];
      phi = fmaf(A.w, __builtin_amdgcn_rsqf(d2), phi);
    }
    if (blocked) continue;
    const float u = fmaf(-k.coul, phi, P.w);
    const float E = fast_exp2(-(u * u) * k.kq);
    acc.V += E;
    if (GRAD) {
      const float c = E * u * k.c2;
      for (int m = 0; m < Ma; ++m) {
        const float4 A = atm_f[m];
        const float ex = px - A.x, ey = py - A.y, ez = pz - A.z;
        const float ri = __builtin_amdgcn_rsqf(fmaf(ex, ex, fmaf(ey, ey, ez * ez)));
        const float w = c * A.w * (ri * ri * ri);
        const float vx = w * ex, vy = w * ey, vz = w * ez;
        f[0] += vx; f[1] += vy; f[2] += vz;
        Fm[0] = fmaf(vx, A.x, Fm[0]); Fm[1] = fmaf(vx, A.y, Fm[1]); Fm[2] = fmaf(vx, A.z, Fm[2]);
        Fm[3] = fmaf(vy, A.x, Fm[3]); Fm[4] = fmaf(vy, A.y, Fm[4]); Fm[5] = fmaf(vy, A.z, Fm[5]);
        Fm[6] = fmaf(vz, A.x, Fm[6]); Fm[7] = fmaf(vz, A.y, Fm[7]); Fm[8] = fmaf(vz, A.z, Fm[8]);
      }
    }
  }
  if (GRAD) {   // fit-frame forces -> world: gt += R f, G += R Fm
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      acc.gt[r] += R[3 * r] * f[0] + R[3 * r + 1] * f[1] + R[3 * r + 2] * f[2];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        acc.G[3 * r + c] += R[3 * r] * Fm[c] + R[3 * r + 1] * Fm[3 + c] + R[3 * r + 2] * Fm[6 + c];
    }
  }
  for (int jj = j; jj < Ms; jj += lpp) {
    const float4 P0 = srf_f[jj];
    const float px = fmaf(R[0], P0.x, fmaf(R[1], P0.y, fmaf(R[2], P0.z, tr[0])));
    const float py = fmaf(R[3], P0.x, fmaf(R[4], P0.y, fmaf(R[5], P0.z, tr[1])));
    const float pz = fmaf(R[6], P0.x, fmaf(R[7], P0.y, fmaf(R[8], P0.z, tr[2])));
    float phi = 0.f;
    bool blocked = false;
    for (int m = 0; m < Na; ++m) {
      const float4 A = atm_r[m];
      const float ex = px - A.x, ey = py - A.y, ez = pz - A.z;
      const float d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
      blocked |= d2 < rb_r[m];
      phi = fmaf(A.w, __builtin_amdgcn_rsqf(d2), phi);
    }
    if (blocked) continue;
    const float u = fmaf(-k.coul, phi, P0.w);
    const float E = fast_exp2(-(u * u) * k.kq);
    acc.V += E;
    if (GRAD) {
      float sx = 0.f, sy = 0.f, sz = 0.f;
      for (int m = 0; m < Na; ++m) {
        const float4 A = atm_r[m];
        const float ex = px - A.x, ey = py - A.y, ez = pz - A.z;
        const float ri = __builtin_amdgcn_rsqf(fmaf(ex, ex, fmaf(ey, ey, ez * ez)));
        const float w = A.w * (ri * ri * ri);
        sx = fmaf(w, ex, sx); sy = fmaf(w, ey, sy); sz = fmaf(w, ez, sz);
      }
      const float c = -E * u * k.c2;
      const float hx = c * sx, hy = c * sy, hz = c * sz;
      acc.gt[0] += hx; acc.gt[1] += hy; acc.gt[2] += hz;
      acc.G[0] = fmaf(hx, P0.x, acc.G[0]); acc.G[1] = fmaf(hx, P0.y, acc.G[1]); acc.G[2] = fmaf(hx, P0.z, acc.G[2]);
      acc.G[3] = fmaf(hy, P0.x, acc.G[3]); acc.G[4] = fmaf(hy, P0.y, acc.G[4]); acc.G[5] = fmaf(hy, P0.z, acc.G[5]);
      acc.G[6] = fmaf(hz, P0.x, acc.G[6]); acc.G[7] = fmaf(hz, P0.y, acc.G[7]); acc.G[8] = fmaf(hz, P0.z, acc.G[8]);
    }
  }
}

// FP64 value of the ShaEP comparison of one pose from the packed global arrays (ref rows in the
// world frame, fit rows in the fit frame, pose x -> R x + t); lanes j of nt.
__device__ __forceinline__ double shaep_value_f64(const float4* __restrict__ srf_r, int Ns,
                                                  const float4* __restrict__ atm_f, const float* __restrict__ rad_f,
                                                  int Ma, const float4* __restrict__ srf_f, int Ms,
                                                  const float4* __restrict__ atm_r, const float* __restrict__ rad_r,
                                                  int Na, int j, int nt, const double R[9], const double t[3],
                                                  double coul, double lam, double probe) {
  double v = 0.0;
  for (int i = j; i < Ns; i += nt) {
    const float4 P = srf_r[i];
    const double dx = P.x - t[0], dy = P.y - t[1], dz = P.z - t[2];
    const double px = R[0] * dx + R[3] * dy + R[6] * dz;
    const double py = R[1] * dx + R[4] * dy + R[7] * dz;
    const double pz = R[2] * dx + R[5] * dy + R[8] * dz;
    double phi = 0.0;
    bool blocked = false;
    for (int m = 0; m < Ma; ++m) {
      const float4 A = atm_f[m];
      const double ex = px - A.x, ey = py - A.y, ez = pz - A.z;
      const double d = sqrt(ex * ex + ey * ey + ez * ez);
      blocked |= d < (double)rad_f[m] + probe;
      phi += (double)A.w / d;
    }
    if (!blocked) {
      const double u = (double)P.w - coul * phi;
      v += exp(-(u * u) / lam);
    }
  }
  for (int jj = j; jj < Ms; jj += nt) {
    const float4 P0 = srf_f[jj];
    const double px = R[0] * P0.x + R[1] * P0.y + R[2] * P0.z + t[0];
    const double py = R[3] * P0.x + R[4] * P0.y + R[5] * P0.z + t[1];
    const double pz = R[6] * P0.x + R[7] * P0.y + R[8] * P0.z + t[2];
    double phi = 0.0;
    bool blocked = false;
    for (int m = 0; m < Na; ++m) {
      const float4 A = atm_r[m];
      const double ex = px - A.x, ey = py - A.y, ez = pz - A.z;
      const double d = sqrt(ex * ex + ey * ey + ez * ez);
      blocked |= d < (double)rad_r[m] + probe;
      phi += (double)A.w / d;
    }
    if (!blocked) {
      const double u = (double)P0.w - coul * phi;
      v += exp(-(u * u) / lam);
    }
  }
  return v;
}

// Convert scaled shape sums into O, dO/dt, dO/dR (in place).
__device__ __forceinline__ void shape_finish(Acc& a, float K, float alpha, float s, bool grad) {
  a.V *= K;
  if (grad) {
    const float ct = K * alpha / s, cr = K * alpha / (s * s);
#pragma unroll
    for (int i = 0; i < 3; ++i) a.gt[i] *= ct;
#pragma unroll
    for (int i = 0; i < 9; ++i) a.G[i] *= cr;
  }
}

// Pharmacophore tables in registers/LDS.
struct PhTab {
  float kap[FSS_N_PTYPES];   // alpha/2 * log2(e)
  float ak[FSS_N_PTYPES];    // alpha * K
  float k[FSS_N_PTYPES];     // K
  int cat[FSS_N_PTYPES];
  int ext;                   // some type has category 4 or 5 (extended points)
};

// ---------------------------------------------------------------------------------------------
// Pharmacophore channel, FP32.  anc[i] = (x, y, z, type bits), vec[i] = (unit x, y, z, 0); the
// ref rows are sorted by type, so fit row b (type t) visits only ref rows [poff[t], poff[t+1]).
// Categories 4 and 5 are the extended-point form of a directional type: the anchor pair term
// has weight 1 (4) or 0 (5) and no vector factor, and a second isotropic term couples the
// extended points anchor + unit vector of the two features.  That term runs as a second pass,
// only when the table has such a category, so the common loop keeps its register budget.  The
// fit extended point moves with the pose as b + u, so its force f_e enters dO/dt as f_e and
// dO/dR as f_e (x) (p + u0).
// ---------------------------------------------------------------------------------------------
template <bool GRAD>
__device__ __forceinline__ void pharm_accum(const float4* __restrict__ ranc,
                                            const float4* __restrict__ rvec,
                                            const int* __restrict__ poff,
                                            const float4* __restrict__ fanc,
                                            const float4* __restrict__ fvec, int M, int j, int lpp,
                                            const float R[9], const float t[3], const PhTab& tab,
                                            Acc& acc) {
  for (int b = j; b < M; b += lpp) {
    const float4 p = fanc[b];
    const float4 u = fvec[b];
    const int tb = __float_as_int(p.w);
    if (tb < 0 || tb >= FSS_N_PTYPES) continue;
    const int cat = tab.cat[tb];
    if (cat == 3) continue;
    const float kap = tab.kap[tb], ak = tab.ak[tb], kk = tab.k[tb];
    const float bx = fmaf(R[0], p.x, fmaf(R[1], p.y, fmaf(R[2], p.z, t[0])));
    const float by = fmaf(R[3], p.x, fmaf(R[4], p.y, fmaf(R[5], p.z, t[1])));
    const float bz = fmaf(R[6], p.x, fmaf(R[7], p.y, fmaf(R[8], p.z, t[2])));
    const float ux = R[0] * u.x + R[1] * u.y + R[2] * u.z;
    const float uy = R[3] * u.x + R[4] * u.y + R[5] * u.z;
    const float uz = R[6] * u.x + R[7] * u.y + R[8] * u.z;
    float fx = 0.f, fy = 0.f, fz = 0.f, hx = 0.f, hy = 0.f, hz = 0.f, v = 0.f;
    const int a_end = poff[tb + 1];
    for (int a = poff[tb]; a < a_end; ++a) {
      const float4 r = ranc[a];
      const float4 rv = rvec[a];
      const float dx = r.x - bx, dy = r.y - by, dz = r.z - bz;
      const float e = fast_exp2(-kap * fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
      const float D = ux * rv.x + uy * rv.y + uz * rv.z;
      float w = 1.f, c = 0.f;
      if (cat == 1) {
        w = (fminf(fmaxf(D, 0.f), 1.f) + 2.f) * (1.f / 3.f);
        c = (D > 0.f && D < 1.f) ? 1.f : 0.f;
      } else if (cat == 2) {
        w = (fabsf(D) + 2.f) * (1.f / 3.f);
        c = D > 0.f ? 1.f : (D < 0.f ? -1.f : 0.f);
      } else if (cat == 5) {
        w = 0.f;
      }
      const float we = w * e;
      v = fmaf(kk, we, v);
      if (GRAD) {
        const float fw = ak * we;
        fx = fmaf(fw, dx, fx); fy = fmaf(fw, dy, fy); fz = fmaf(fw, dz, fz);
        const float hc = kk * (1.f / 3.f) * e * c;
        hx = fmaf(hc, rv.x, hx); hy = fmaf(hc, rv.y, hy); hz = fmaf(hc, rv.z, hz);
      }
    }
    acc.V += v;
    if (GRAD) {
      acc.gt[0] += fx; acc.gt[1] += fy; acc.gt[2] += fz;
      acc.G[0] += fx * p.x + hx * u.x; acc.G[1] += fx * p.y + hx * u.y; acc.G[2] += fx * p.z + hx * u.z;
      acc.G[3] += fy * p.x + hy * u.x; acc.G[4] += fy * p.y + hy * u.y; acc.G[5] += fy * p.z + hy * u.z;
      acc.G[6] += fz * p.x + hz * u.x; acc.G[7] += fz * p.y + hz * u.y; acc.G[8] += fz * p.z + hz * u.z;
    }
  }
  if (!tab.ext) return;
  for (int b = j; b < M; b += lpp) {   // extended points of categories 4 and 5
    const float4 p = fanc[b];
    const int tb = __float_as_int(p.w);
    if (tb < 0 || tb >= FSS_N_PTYPES || tab.cat[tb] < 4) continue;
    const float4 u = fvec[b];
    const float kap = tab.kap[tb], ak = tab.ak[tb], kk = tab.k[tb];
    // the moved extended point R (p + u) + t
    const float px = p.x + u.x, py = p.y + u.y, pz = p.z + u.z;
    const float ex0 = fmaf(R[0], px, fmaf(R[1], py, fmaf(R[2], pz, t[0])));
    const float ey0 = fmaf(R[3], px, fmaf(R[4], py, fmaf(R[5], pz, t[1])));
    const float ez0 = fmaf(R[6], px, fmaf(R[7], py, fmaf(R[8], pz, t[2])));
    float fx = 0.f, fy = 0.f, fz = 0.f, v = 0.f;
    const int a_end = poff[tb + 1];
    for (int a = poff[tb]; a < a_end; ++a) {
      const float4 r = ranc[a];
      const float4 rv = rvec[a];
      const float ex = (r.x + rv.x) - ex0, ey = (r.y + rv.y) - ey0, ez = (r.z + rv.z) - ez0;
      const float e2 = fast_exp2(-kap * fmaf(ex, ex, fmaf(ey, ey, ez * ez)));
      v = fmaf(kk, e2, v);
      if (GRAD) {
        const float fe = ak * e2;
        fx = fmaf(fe, ex, fx); fy = fmaf(fe, ey, fy); fz = fmaf(fe, ez, fz);
      }
    }
    acc.V += v;
    if (GRAD) {
      acc.gt[0] += fx; acc.gt[1] += fy; acc.gt[2] += fz;
      acc.G[0] += fx * px; acc.G[1] += fx * py; acc.G[2] += fx * pz;
      acc.G[3] += fy * px; acc.G[4] += fy * py; acc.G[5] += fy * pz;
      acc.G[6] += fz * px; acc.G[7] += fz * py; acc.G[8] += fz * pz;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// FP64 value-only evaluation straight from the packed (unscaled) device arrays; used for the
// final re-scoring of every optimised pose.
// ---------------------------------------------------------------------------------------------
template <bool ESP>
__device__ __forceinline__ double shape_value_f64(const float4* __restrict__ ref, int N,
                                                  const float4* __restrict__ fit, int M, int j,
                                                  int lpp, const double R[9], const double t[3],
                                                  double half_alpha, double inv_lam) {
  double v = 0.0;
  for (int b = j; b < M; b += lpp) {
    const float4 p = fit[b];
    const double bx = R[0] * p.x + R[1] * p.y + R[2] * p.z + t[0];
    const double by = R[3] * p.x + R[4] * p.y + R[5] * p.z + t[1];
    const double bz = R[6] * p.x + R[7] * p.y + R[8] * p.z + t[2];
    for (int a = 0; a < N; ++a) {
      const float4 r = ref[a];
      const double dx = r.x - bx, dy = r.y - by, dz = r.z - bz;
      double arg = -half_alpha * (dx * dx + dy * dy + dz * dz);
      if (ESP) {
        const double dc = (double)r.w - (double)p.w;
        arg -= dc * dc * inv_lam;
      }
      v += exp(arg);
    }
  }
  return v;
}

// FP32 counterpart of shape_value_f64 (direct form, exp2 with the scale folded in: c_xyz =
// -alpha/2 log2 e, c_w = -log2 e / lam) for the re-scoring pre-pass that picks the restarts
// worth an FP64 score.
template <bool ESP>
__device__ __forceinline__ float shape_value_f32(const float4* __restrict__ ref, int N,
                                                 const float4* __restrict__ fit, int M, int j, int lpp,
                                                 const float R[9], const float t[3], float c_xyz, float c_w) {
  float v0 = 0.f, v1 = 0.f;
  for (int b = j; b < M; b += lpp) {
    const float4 p = fit[b];
    const float bx = fmaf(R[0], p.x, fmaf(R[1], p.y, fmaf(R[2], p.z, t[0])));
    const float by = fmaf(R[3], p.x, fmaf(R[4], p.y, fmaf(R[5], p.z, t[1])));
    const float bz = fmaf(R[6], p.x, fmaf(R[7], p.y, fmaf(R[8], p.z, t[2])));
    int a = 0;
    for (; a + 1 < N; a += 2) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const float4 r = ref[a + k];
        const float dx = r.x - bx, dy = r.y - by, dz = r.z - bz;
        float arg = c_xyz * fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (ESP) { const float dc = r.w - p.w; arg = fmaf(c_w * dc, dc, arg); }
        (k ? v1 : v0) += fast_exp2(arg);
      }
    }
    if (a < N) {
      const float4 r = ref[a];
      const float dx = r.x - bx, dy = r.y - by, dz = r.z - bz;
      float arg = c_xyz * fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      if (ESP) { const float dc = r.w - p.w; arg = fmaf(c_w * dc, dc, arg); }
      v0 += fast_exp2(arg);
    }
  }
  return v0 + v1;
}

__device__ __forceinline__ double pharm_value_f64(const float4* __restrict__ ranc,
                                                  const float4* __restrict__ rvec,
                                                  const int32_t* __restrict__ rtyp, int N,
                                                  const float4* __restrict__ fanc,
                                                  const float4* __restrict__ fvec,
                                                  const int32_t* __restrict__ ftyp, int M, int j,
                                                  int lpp, const double R[9], const double t[3],
                                                  const double* alpha, const double* kk,
                                                  const int32_t* cats) {
  double v = 0.0;
  for (int b = j; b < M; b += lpp) {
    const int tb = ftyp[b];
    if (tb < 0 || tb >= FSS_N_PTYPES) continue;
    const int cat = cats[tb];
    if (cat == 3) continue;
    const float4 p = fanc[b];
    const float4 u0 = fvec[b];
    double un = sqrt((double)u0.x * u0.x + (double)u0.y * u0.y + (double)u0.z * u0.z);
    un = un > 1e-12 ? un : 1e-12;
    const double ux0 = u0.x / un, uy0 = u0.y / un, uz0 = u0.z / un;
    const double bx = R[0] * p.x + R[1] * p.y + R[2] * p.z + t[0];
    const double by = R[3] * p.x + R[4] * p.y + R[5] * p.z + t[1];
    const double bz = R[6] * p.x + R[7] * p.y + R[8] * p.z + t[2];
    const double ux = R[0] * ux0 + R[1] * uy0 + R[2] * uz0;
    const double uy = R[3] * ux0 + R[4] * uy0 + R[5] * uz0;
    const double uz = R[6] * ux0 + R[7] * uy0 + R[8] * uz0;
    const double ha = 0.5 * alpha[tb], K = kk[tb];
    for (int a = 0; a < N; ++a) {
      if (rtyp[a] != tb) continue;
      const float4 r = ranc[a];
      const float4 rv0 = rvec[a];
      double rn = sqrt((double)rv0.x * rv0.x + (double)rv0.y * rv0.y + (double)rv0.z * rv0.z);
      rn = rn > 1e-12 ? rn : 1e-12;
      const double dx = r.x - bx, dy = r.y - by, dz = r.z - bz;
      const double e = exp(-ha * (dx * dx + dy * dy + dz * dz));
      const double D = (ux * rv0.x + uy * rv0.y + uz * rv0.z) / rn;
      double w = 1.0;
      if (cat == 1) w = (fmin(fmax(D, 0.0), 1.0) + 2.0) / 3.0;
      else if (cat == 2) w = (fabs(D) + 2.0) / 3.0;
      else if (cat == 5) w = 0.0;
      v += K * w * e;
      if (cat >= 4) {   // extended points anchor + unit vector, isotropic
        const double ex = (r.x + rv0.x / rn) - (bx + ux), ey = (r.y + rv0.y / rn) - (by + uy),
                     ez = (r.z + rv0.z / rn) - (bz + uz);
        v += K * exp(-ha * (ex * ex + ey * ey + ez * ez));
      }
    }
  }
  return v;
}

// FP32 counterpart of pharm_value_f64 (same rows, weights and extended points; exp2 with the
// table's alpha/2 log2 e) for the re-scoring pre-pass that picks the restarts worth an FP64 score.
__device__ __forceinline__ float pharm_value_f32(const float4* __restrict__ ranc,
                                                 const float4* __restrict__ rvec,
                                                 const int32_t* __restrict__ rtyp, int N,
                                                 const float4* __restrict__ fanc,
                                                 const float4* __restrict__ fvec,
                                                 const int32_t* __restrict__ ftyp, int M, int j,
                                                 int lpp, const float R[9], const float t[3],
                                                 const PhTab& tab) {
  float v = 0.f;
  for (int b = j; b < M; b += lpp) {
    const int tb = ftyp[b];
    if (tb < 0 || tb >= FSS_N_PTYPES) continue;
    const int cat = tab.cat[tb];
    if (cat == 3) continue;
    const float4 p = fanc[b];
    const float4 u0 = fvec[b];
    const float un = fmaxf(sqrtf(u0.x * u0.x + u0.y * u0.y + u0.z * u0.z), 1e-12f);
    const float ux0 = u0.x / un, uy0 = u0.y / un, uz0 = u0.z / un;
    const float bx = fmaf(R[0], p.x, fmaf(R[1], p.y, fmaf(R[2], p.z, t[0])));
    const float by = fmaf(R[3], p.x, fmaf(R[4], p.y, fmaf(R[5], p.z, t[1])));
    const float bz = fmaf(R[6], p.x, fmaf(R[7], p.y, fmaf(R[8], p.z, t[2])));
    const float ux = R[0] * ux0 + R[1] * uy0 + R[2] * uz0;
    const float uy = R[3] * ux0 + R[4] * uy0 + R[5] * uz0;
    const float uz = R[6] * ux0 + R[7] * uy0 + R[8] * uz0;
    const float kap = tab.kap[tb], kk = tab.k[tb];
    for (int a = 0; a < N; ++a) {
      if (rtyp[a] != tb) continue;
      const float4 r = ranc[a];
      const float4 rv0 = rvec[a];
      const float rn = fmaxf(sqrtf(rv0.x * rv0.x + rv0.y * rv0.y + rv0.z * rv0.z), 1e-12f);
      const float dx = r.x - bx, dy = r.y - by, dz = r.z - bz;
      const float e = fast_exp2(-kap * fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
      const float D = (ux * rv0.x + uy * rv0.y + uz * rv0.z) / rn;
      float w = 1.f;
      if (cat == 1) w = (fminf(fmaxf(D, 0.f), 1.f) + 2.f) * (1.f / 3.f);
      else if (cat == 2) w = (fabsf(D) + 2.f) * (1.f / 3.f);
      else if (cat == 5) w = 0.f;
      v = fmaf(kk * w, e, v);
      if (cat >= 4) {   // extended points anchor + unit vector, isotropic
        const float ex = (r.x + rv0.x / rn) - (bx + ux), ey = (r.y + rv0.y / rn) - (by + uy),
                    ez = (r.z + rv0.z / rn) - (bz + uz);
        v = fmaf(kk, fast_exp2(-kap * fmaf(ex, ex, fmaf(ey, ey, ez * ez))), v);
      }
    }
  }
  return v;
}

}  // namespace fss
