// libfss_hip.so: gfx950 kernels and the C ABI declared in include/fss.h.
//
// Kernels
//   k_self_overlap   one workgroup per (pair, side): VAA / VBB in FP64 at the identity pose
//   k_pose_grad      one workgroup per pose: O_AB + dO/dR, dO/dt, dO/dq_hat (per channel), or the
//                    full per-pose loss and SE(3) gradient of one optimiser step
//   k_adam           torch.optim.Adam step over (n, 7) parameter rows
//   k_adam_qt        the accelerated drivers' Adam on unit quaternions (no bias correction)
//   k_align          persistent optimiser: one workgroup per (pair, restart-block); the pair's
//                    clouds are staged once in LDS, every lane group of `lpp` lanes owns one
//                    restart for all steps (parameters and Adam moments live in registers), the
//                    N x M sum of a pose is split over the group and reduced with butterflies,
//                    and per-pair early stop is one LDS exchange per step
//   k_align_finish   best restart per pair + SE(3) matrix
//   k_score_fwd/bwd  differentiable similarity of fixed clouds (FP32 and FP64 instances)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/fss.h"
#include "fss_common.h"
#include "fss_device.h"

using namespace fss;

namespace {

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error((int)e, "%s: launch failed: %s", what, hipGetErrorString(e));
  return FSS_OK;
}

constexpr int kPoseBlock = 256;  // threads of the one-workgroup-per-pose kernels

bool uses_shape(int mode) {
  return mode == FSS_MODE_SHAPE || mode == FSS_MODE_SHAPE_ESP || mode == FSS_MODE_SHAPE_PHARM || mode == FSS_MODE_ESP_COMBO;
}
bool uses_shaep(int mode) { return mode == FSS_MODE_ESP_COMBO; }
bool uses_pharm(int mode) { return mode == FSS_MODE_PHARM || mode == FSS_MODE_SHAPE_PHARM; }

// Device-side scoring constants derived once per call on the host.
struct KScore {
  int mode;
  float alpha, K, s, sc;      // shape: K = (pi/(2 alpha))^1.5, s = sqrt(alpha/2 log2 e), sc = sqrt(log2 e / lam)
  double alpha_d, lam_d;
  float w;
  int sim;
  float sigma;
  double w_d, sigma_d;
  PhTab tab;
  double ph_alpha[FSS_N_PTYPES];
  double ph_k[FSS_N_PTYPES];
  int32_t ph_cat[FSS_N_PTYPES];
  // ShaEP channel (FSS_MODE_ESP_COMBO)
  ShaepK se;
  float se_w, se_probe2;         // esp weight, probe radius (staging adds it to each radius)
  double se_w_d, se_coul_d, se_lam_d, se_probe_d;
};

int make_kscore(const FssScoreParams* sp, KScore& ks) {
  if (!sp) return set_error(FSS_E_INVALID, "score params are NULL");
  if (sp->mode < 0 || sp->mode > 4) return set_error(FSS_E_INVALID, "unknown mode %d", sp->mode);
  if (uses_shape(sp->mode) && !(sp->alpha > 0.0)) return set_error(FSS_E_INVALID, "alpha must be > 0");
  if (sp->mode == FSS_MODE_ESP_COMBO) {
    if (!(sp->lam > 0.0)) return set_error(FSS_E_INVALID, "lam must be > 0");
    if (!(sp->coulomb > 0.0)) return set_error(FSS_E_INVALID, "coulomb must be > 0");
    if (!(sp->probe_radius >= 0.0)) return set_error(FSS_E_INVALID, "probe_radius must be >= 0");
    if (!(sp->esp_weight >= 0.0 && sp->esp_weight <= 1.0)) return set_error(FSS_E_INVALID, "esp_weight must be in [0, 1]");
  }
  if (sp->mode == FSS_MODE_SHAPE_ESP && !(sp->lam > 0.0)) return set_error(FSS_E_INVALID, "lam must be > 0");
  if (sp->similarity != FSS_SIM_TANIMOTO && sp->similarity != FSS_SIM_TVERSKY)
    return set_error(FSS_E_INVALID, "unknown similarity %d", sp->similarity);
  if (uses_pharm(sp->mode))
    for (int i = 0; i < FSS_N_PTYPES; ++i)
      if (sp->ph_cat[i] < 0 || sp->ph_cat[i] > 5)
        return set_error(FSS_E_INVALID, "ph_cat[%d] = %d is not a category (0-5)", i, sp->ph_cat[i]);
  memset(&ks, 0, sizeof(ks));
  ks.mode = sp->mode;
  ks.alpha = (float)sp->alpha;
  ks.alpha_d = sp->alpha;
  ks.K = uses_shape(sp->mode) ? (float)pow(M_PI / (2.0 * sp->alpha), 1.5) : 0.f;
  ks.s = uses_shape(sp->mode) ? (float)sqrt(0.5 * sp->alpha * kLog2eD) : 1.f;
  ks.lam_d = sp->lam > 0.0 ? sp->lam : 1.0;
  ks.sc = sp->mode == FSS_MODE_SHAPE_ESP ? (float)sqrt(kLog2eD / sp->lam) : 0.f;
  ks.w = (float)sp->color_weight;
  ks.w_d = sp->color_weight;
  ks.sim = sp->similarity;
  ks.sigma = (float)sp->sigma;
  ks.sigma_d = sp->sigma;
  for (int i = 0; i < FSS_N_PTYPES; ++i) {
    ks.ph_alpha[i] = sp->ph_alpha[i];
    ks.ph_k[i] = sp->ph_k[i];
    ks.ph_cat[i] = sp->ph_cat[i];
    ks.tab.kap[i] = (float)(0.5 * sp->ph_alpha[i] * kLog2eD);
    ks.tab.ak[i] = (float)(sp->ph_alpha[i] * sp->ph_k[i]);
    ks.tab.k[i] = (float)sp->ph_k[i];
    ks.tab.cat[i] = sp->ph_cat[i];
    ks.tab.ext |= sp->ph_cat[i] >= 4;
  }
  if (sp->mode == FSS_MODE_ESP_COMBO) {
    ks.lam_d = 1.0;                    // the shape channel has no ESP weighting here
    ks.se.coul = (float)sp->coulomb;
    ks.se.kq = (float)(kLog2eD / sp->lam);
    ks.se.c2 = (float)(2.0 * sp->coulomb / sp->lam);
    ks.se_w = (float)sp->esp_weight;
    ks.se_probe2 = (float)sp->probe_radius;
    ks.se_w_d = sp->esp_weight;
    ks.se_coul_d = sp->coulomb;
    ks.se_lam_d = sp->lam;
    ks.se_probe_d = sp->probe_radius;
  }
  return FSS_OK;
}

int check_batch(const FssPairBatch* pb, int mode) {
  if (!pb) return set_error(FSS_E_INVALID, "pair batch is NULL");
  if (pb->n_pairs < 0) return set_error(FSS_E_INVALID, "n_pairs < 0");
  if (uses_shape(mode) && (!pb->ref_pts || !pb->fit_pts || !pb->ref_rng || !pb->fit_rng))
    return set_error(FSS_E_INVALID, "shape channel arrays are NULL");
  if (uses_pharm(mode) && (!pb->ref_anc || !pb->ref_vec || !pb->ref_typ || !pb->ref_prng ||
                           !pb->fit_anc || !pb->fit_vec || !pb->fit_typ || !pb->fit_prng))
    return set_error(FSS_E_INVALID, "pharmacophore channel arrays are NULL");
  if (pb->max_ref < 0 || pb->max_fit < 0 || pb->max_ref_ph < 0 || pb->max_fit_ph < 0)
    return set_error(FSS_E_INVALID, "negative cloud maxima");
  if (uses_shaep(mode)) {
    if (!pb->ref_srf || !pb->fit_srf || !pb->ref_srng || !pb->fit_srng || !pb->ref_atm || !pb->fit_atm ||
        !pb->ref_rad || !pb->fit_rad || !pb->ref_arng || !pb->fit_arng)
      return set_error(FSS_E_INVALID, "ShaEP channel arrays are NULL");
    if (pb->max_ref_srf < 0 || pb->max_fit_srf < 0 || pb->max_ref_atm < 0 || pb->max_fit_atm < 0)
      return set_error(FSS_E_INVALID, "negative ShaEP cloud maxima");
  }
  return FSS_OK;
}

// LDS plan for one pair (float4 offsets): shape ref rows, fit rows, ref ESP column, one float4
// for the ref centroid, then pharm ref anchors / vectors, fit anchors / vectors.
// With `mfma` (inner-loop variant 3) the ref rows are the split-bf16 A-operand tiles (64 float4
// per 16 rows) followed by the 2 a~ row quads (12 float4 per 16 rows), and no ESP column.
struct LdsPlan {
  int ref, xq, fit, refc, red, poff, ranc, rvec, fanc, fvec;
  int srf_r, srf_f, atm_r, atm_f, rb_r, rb_f;   // ShaEP channel
  int end;
  int tile_ref, tile_fit;   // > 0: streamed shape clouds, this many rows of each side per tile
};
LdsPlan plan_lds(const FssPairBatch* pb, int mode, bool mfma = false) {
  LdsPlan p;
  int o = 0;
  const bool sh = uses_shape(mode), ph = uses_pharm(mode);
  const int tiles = (pb->max_ref + 15) / 16;
  p.ref = o; o += sh ? (mfma ? 64 * tiles : pb->max_ref) : 0;
  p.xq = o; o += sh && mfma ? 12 * tiles : 0;
  p.fit = o; o += sh ? pb->max_fit : 0;
  p.refc = o; o += mode == FSS_MODE_SHAPE_ESP && !mfma ? (pb->max_ref + 3) / 4 : 0;
  p.red = o; o += 1;                 // the ref centroid of stage_pair (3 floats)
  p.poff = o; o += ph ? 4 : 0;
  p.ranc = o; o += ph ? pb->max_ref_ph : 0;
  p.rvec = o; o += ph ? pb->max_ref_ph : 0;
  p.fanc = o; o += ph ? pb->max_fit_ph : 0;
  p.fvec = o; o += ph ? pb->max_fit_ph : 0;
  const bool se = uses_shaep(mode);
  p.srf_r = o; o += se ? pb->max_ref_srf : 0;
  p.srf_f = o; o += se ? pb->max_fit_srf : 0;
  p.atm_r = o; o += se ? pb->max_ref_atm : 0;
  p.atm_f = o; o += se ? pb->max_fit_atm : 0;
  p.rb_r = o; o += se ? (pb->max_ref_atm + 3) / 4 : 0;
  p.rb_f = o; o += se ? (pb->max_fit_atm + 3) / 4 : 0;
  p.end = o;
  p.tile_ref = p.tile_fit = 0;
  return p;
}

// Modes whose shape clouds may be streamed through LDS tiles (clouds of any size); the
// pharmacophore rows of the shape + pharmacophore mode stay staged whole.
constexpr bool tileable(int mode) {
  return mode == FSS_MODE_SHAPE || mode == FSS_MODE_SHAPE_ESP || mode == FSS_MODE_SHAPE_PHARM;
}

// LDS plan of the streamed shape path: one tile of tr ref rows (expanded form, plus the ESP
// column) and one of tf fit rows, the ref centroid slot, then the pair's pharmacophore rows
// (shape + pharmacophore mode) as plan_lds lays them out.
LdsPlan plan_lds_tiled(const FssPairBatch* pb, int mode, int tr, int tf) {
  LdsPlan p{};
  int o = 0;
  const bool ph = mode == FSS_MODE_SHAPE_PHARM;
  p.ref = o; o += tr;
  p.xq = o;
  p.fit = o; o += tf;
  p.refc = o; o += mode == FSS_MODE_SHAPE_ESP ? (tr + 3) / 4 : 0;
  p.red = o; o += 1;
  p.poff = o; o += ph ? 4 : 0;
  p.ranc = o; o += ph ? pb->max_ref_ph : 0;
  p.rvec = o; o += ph ? pb->max_ref_ph : 0;
  p.fanc = o; o += ph ? pb->max_fit_ph : 0;
  p.fvec = o; o += ph ? pb->max_fit_ph : 0;
  p.srf_r = p.srf_f = p.atm_r = p.atm_f = p.rb_r = p.rb_f = o;
  p.end = o;
  p.tile_ref = tr;
  p.tile_fit = tf;
  return p;
}

constexpr size_t kCuLdsBytes = 160 * 1024;   // LDS of one gfx950 compute unit

// LDS per workgroup on the current device; without one (host-only planning) gfx950's 160 KiB.
size_t max_lds_bytes() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kCuLdsBytes;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return kCuLdsBytes;
  return (size_t)v;
}

struct Staged {
  const float4* ref; const float* refc; const float4* fit; int N, M;
  const float4* xq; int T;  // mfma layout: ref = A-operand tiles, xq = 2 a~ row quads, T tiles
  float cs[3];  // scaled ref centroid: the origin of the expanded shape form
  const float4* ranc; const float4* rvec; const float4* fanc; const float4* fvec; int Np, Mp;
  const int* poff;  // ref pharmacophore rows are sorted by type; type t occupies [poff[t], poff[t+1])
  // ShaEP channel: ref rows relative to the origin co (the ref centroid), fit rows raw
  const float4* srf_r; const float4* srf_f; const float4* atm_r; const float4* atm_f;
  const float* rb_r; const float* rb_f;   // (radius + probe)^2
  int Ns, Ms, Na, Ma;
  float co[3];
  const float4* gref; const float4* gfit;   // streamed shape path: the pair's rows in HBM
};

// Pharmacophore rows of one pair into LDS (anchor with the type in w as integer bits, unit
// vector): ref rows in type-sorted order with the per-type offsets, fit rows as given.  Caller
// syncs afterwards.  ref_side_a / fit_side_b as in stage_pair.
__device__ __forceinline__ void stage_pharm(const FssPairBatch& pb, LdsPlan lp, int pair, int ref_side_a,
                                            int fit_side_b, float4* lds, Staged& st) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int* ra = ref_side_a ? pb.fit_prng : pb.ref_prng;
  const int* rb = fit_side_b ? pb.fit_prng : pb.ref_prng;
  const float4* aa = reinterpret_cast<const float4*>(ref_side_a ? pb.fit_anc : pb.ref_anc);
  const float4* av = reinterpret_cast<const float4*>(ref_side_a ? pb.fit_vec : pb.ref_vec);
  const int32_t* at = ref_side_a ? pb.fit_typ : pb.ref_typ;
  const float4* ba = reinterpret_cast<const float4*>(fit_side_b ? pb.fit_anc : pb.ref_anc);
  const float4* bv = reinterpret_cast<const float4*>(fit_side_b ? pb.fit_vec : pb.ref_vec);
  const int32_t* bt = fit_side_b ? pb.fit_typ : pb.ref_typ;
  const int a0 = ra[2 * pair], N = ra[2 * pair + 1];
  const int b0 = rb[2 * pair], M = rb[2 * pair + 1];
  // Ref rows go to their type-sorted position (a stable counting sort by direct ranking:
  // the clouds hold a few dozen features); out-of-range types are filed as Dummy.
  auto tclamp = [](int ty) { return (ty < 0 || ty >= FSS_N_PTYPES) ? FSS_N_PTYPES - 1 : ty; };
  int* poff = reinterpret_cast<int*>(lds + lp.poff);
  if (tid <= FSS_N_PTYPES) {
    int c = 0;
    for (int a = 0; a < N; ++a) c += tclamp(at[a0 + a]) < tid;
    poff[tid] = c;
  }
  for (int i = tid; i < N + M; i += nt) {
    const bool isref = i < N;
    const int k = isref ? a0 + i : b0 + (i - N);
    const float4 p = isref ? aa[k] : ba[k];
    const float4 v = isref ? av[k] : bv[k];
    const int ty = tclamp(isref ? at[k] : bt[k]);
    float n = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    n = n > 1e-12f ? n : 1e-12f;
    const float4 pa4 = make_float4(p.x, p.y, p.z, __int_as_float(ty));
    const float4 va4 = make_float4(v.x / n, v.y / n, v.z / n, 0.f);
    if (isref) {
      int pos = 0;
      for (int a = 0; a < N; ++a) {
        const int ta = tclamp(at[a0 + a]);
        pos += (ta < ty) || (ta == ty && a < i);
      }
      lds[lp.ranc + pos] = pa4;
      lds[lp.rvec + pos] = va4;
    } else {
      lds[lp.fanc + (i - N)] = pa4;
      lds[lp.fvec + (i - N)] = va4;
    }
  }
  st.ranc = lds + lp.ranc; st.rvec = lds + lp.rvec; st.fanc = lds + lp.fanc; st.fvec = lds + lp.fvec;
  st.Np = N; st.Mp = M; st.poff = poff;
}

// Ref centroid of rows ga[0..N): lane-strided sums in wavefront 0 and one butterfly, whatever the
// workgroup size, so that every tiling and every kernel staging this pair gets the same bits.
// Every thread of the workgroup calls it (one barrier inside); red: 3 floats of LDS.
__device__ __forceinline__ void ref_centroid(const float4* __restrict__ ga, int N, float* red, float c[3]) {
  const int tid = threadIdx.x;
  float cx = 0.f, cy = 0.f, cz = 0.f;
  if (tid < 64) {
    for (int i = tid; i < N; i += 64) {
      const float4 v = ga[i];
      cx += v.x; cy += v.y; cz += v.z;
    }
    cx = group_sum(cx, 64); cy = group_sum(cy, 64); cz = group_sum(cz, 64);
    if (tid == 0) { red[0] = cx; red[1] = cy; red[2] = cz; }
  }
  __syncthreads();
  const float invN = N > 0 ? 1.f / (float)N : 0.f;
  c[0] = red[0] * invN; c[1] = red[1] * invN; c[2] = red[2] * invN;
}

// Expanded-form LDS rows of ref rows ga[i0 .. i0+n) (see shape_accum): la[i] = (2 a~, W_a - D),
// lc[i] = 2 c~_a (ESP only), a~ = s (a - c).
template <bool ESP>
__device__ __forceinline__ void stage_ref_rows(const float4* __restrict__ ga, int i0, int n, const float c[3],
                                               float s, float sc, float4* la, float* lc) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float4 v = ga[i0 + i];
    const float ax = s * (v.x - c[0]), ay = s * (v.y - c[1]), az = s * (v.z - c[2]);
    const float cc = ESP ? v.w * sc : 0.f;
    la[i] = make_float4(2.f * ax, 2.f * ay, 2.f * az, -fmaf(ax, ax, fmaf(ay, ay, fmaf(az, az, cc * cc))) - kExpShift);
    if (ESP) lc[i] = 2.f * cc;
  }
}

// Scaled fit rows gb[i0 .. i0+n): lb[i] = (s b, c~_b).
template <bool ESP>
__device__ __forceinline__ void stage_fit_rows(const float4* __restrict__ gb, int i0, int n, float s, float sc,
                                               float4* lb) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float4 v = gb[i0 + i];
    lb[i] = make_float4(v.x * s, v.y * s, v.z * s, ESP ? v.w * sc : 0.f);
  }
}

// Streamed shape path: the pair's centroid and row pointers (the tiles are staged per pass),
// and its pharmacophore rows staged whole (PH).  Caller syncs afterwards.
template <bool ESP, bool PH>
__device__ __forceinline__ Staged stage_tiled(const FssPairBatch& pb, const KScore& ks, LdsPlan lp, int pair,
                                              float4* lds) {
  Staged st{};
  const int a0 = pb.ref_rng[2 * pair], N = pb.ref_rng[2 * pair + 1];
  const int b0 = pb.fit_rng[2 * pair], M = pb.fit_rng[2 * pair + 1];
  st.gref = reinterpret_cast<const float4*>(pb.ref_pts) + a0;
  st.gfit = reinterpret_cast<const float4*>(pb.fit_pts) + b0;
  float c[3];
  ref_centroid(st.gref, N, reinterpret_cast<float*>(lds + lp.red), c);
  st.ref = lds + lp.ref;
  st.refc = reinterpret_cast<const float*>(lds + lp.refc);
  st.fit = lds + lp.fit;
  st.N = N; st.M = M;
  st.cs[0] = ks.s * c[0]; st.cs[1] = ks.s * c[1]; st.cs[2] = ks.s * c[2];
  st.co[0] = c[0]; st.co[1] = c[1]; st.co[2] = c[2];
  if (PH) stage_pharm(pb, lp, pair, 0, 1, lds, st);
  return st;
}

// Tile residency of the streamed shape path (uniform across the workgroup).
struct TileState {
  int r0 = -1, f0 = -1;
};

// One pose's shape sums over every (ref tile, fit tile) of the streamed path.  Every thread of the
// workgroup calls it: a tile is staged between two barriers when it is not already resident;
// `work` lanes accumulate (lane j of a group of lpp lanes takes the tile's fit rows j, j+lpp, ...).
template <bool ESP, bool GRAD>
__device__ __forceinline__ void shape_accum_tiled(const Staged& st, const KScore& ks, LdsPlan lp, TileState& ts_,
                                                  bool work, int j, int lpp, const float R[9], const float ts[3],
                                                  Acc& acc) {
  float4* la = const_cast<float4*>(st.ref);
  float* lc = const_cast<float*>(st.refc);
  float4* lb = const_cast<float4*>(st.fit);
  for (int f0 = 0; f0 < st.M; f0 += lp.tile_fit) {
    const int nf = min(lp.tile_fit, st.M - f0);
    for (int r0 = 0; r0 < st.N; r0 += lp.tile_ref) {
      const int nr = min(lp.tile_ref, st.N - r0);
      if (ts_.r0 != r0 || ts_.f0 != f0) {
        __syncthreads();   // every lane is done with the resident tiles
        if (ts_.r0 != r0) stage_ref_rows<ESP>(st.gref, r0, nr, st.co, ks.s, ks.sc, la, lc);
        if (ts_.f0 != f0) stage_fit_rows<ESP>(st.gfit, f0, nf, ks.s, ks.sc, lb);
        ts_.r0 = r0;
        ts_.f0 = f0;
        __syncthreads();
      }
      if (work) shape_accum<ESP, GRAD>(la, lc, nr, lb, nf, j, lpp, R, ts, acc);
    }
  }
}

// Cooperative copy of one pair's clouds into LDS (shape rows scaled, pharm vectors normalised,
// pharm type folded into the anchor's w as integer bits).  Caller syncs afterwards.
template <bool SH, bool ESP, bool PH, bool EXPD = true, bool MF = false, bool SE = false>
__device__ __forceinline__ Staged stage_pair(const FssPairBatch& pb, const KScore& ks, LdsPlan lp,
                                             int pair, int ref_side_a, int fit_side_b,
                                             float4* lds) {
  // ref_side_a / fit_side_b select which cloud feeds the "ref" and "fit" slots:
  // 0 = the pair's ref cloud, 1 = the pair's fit cloud (self-overlaps use (0,0) and (1,1)).
  Staged st{};
  const int tid = threadIdx.x, nt = blockDim.x;
  if (SH) {
    const int* ra = ref_side_a ? pb.fit_rng : pb.ref_rng;
    const float* pa = ref_side_a ? pb.fit_pts : pb.ref_pts;
    const int* rb = fit_side_b ? pb.fit_rng : pb.ref_rng;
    const float* pbp = fit_side_b ? pb.fit_pts : pb.ref_pts;
    const int a0 = ra[2 * pair], N = ra[2 * pair + 1];
    const int b0 = rb[2 * pair], M = rb[2 * pair + 1];
    const float4* ga = reinterpret_cast<const float4*>(pa) + a0;
    const float4* gb = reinterpret_cast<const float4*>(pbp) + b0;
    float4* la = lds + lp.ref;
    float4* lb = lds + lp.fit;
    float* lc = reinterpret_cast<float*>(lds + lp.refc);
    float* red = reinterpret_cast<float*>(lds + lp.red);
    const float s = ks.s, sc = ks.sc;
    float cen[3];
    ref_centroid(ga, N, red, cen);
    const float cx = cen[0], cy = cen[1], cz = cen[2];
    if (MF) {
      const int T = (N + 15) >> 4;
      float* xqf = reinterpret_cast<float*>(lds + lp.xq);
      for (int i = tid; i < 16 * T; i += nt) {
        float ax = 0.f, ay = 0.f, az = 0.f, cc = 0.f, w = -1e30f;
        if (i < N) {
          const float4 v = ga[i];
          ax = s * (v.x - cx); ay = s * (v.y - cy); az = s * (v.z - cz);
          cc = ESP ? v.w * sc : 0.f;
          w = -fmaf(ax, ax, fmaf(ay, ay, fmaf(az, az, cc * cc))) - kExpShift;
        }
        const int t = i >> 4, r = i & 15;
        float4* at = la + 64 * t + r;
        at[0] = __builtin_bit_cast(float4, mfma_a_coord(2.f * ax));
        at[16] = __builtin_bit_cast(float4, mfma_a_coord(2.f * ay));
        at[32] = __builtin_bit_cast(float4, mfma_a_coord(2.f * az));
        at[48] = __builtin_bit_cast(float4, mfma_a_last(2.f * cc, w));
        float* xr = xqf + 48 * t + 12 * (r >> 2) + (r & 3);
        xr[0] = 2.f * ax; xr[4] = 2.f * ay; xr[8] = 2.f * az;
      }
      st.xq = lds + lp.xq;
      st.T = T;
    }
    if (EXPD && !MF) {
      stage_ref_rows<ESP>(ga, 0, N, cen, s, sc, la, lc);
    } else {
      for (int i = tid; i < N && !MF; i += nt) {
        const float4 v = ga[i];
        const float ax = s * (v.x - cx), ay = s * (v.y - cy), az = s * (v.z - cz);
        la[i] = make_float4(ax, ay, az, ESP ? v.w * sc : 0.f);
      }
    }
    stage_fit_rows<ESP>(gb, 0, M, s, sc, lb);
    st.ref = la; st.refc = lc; st.fit = lb; st.N = N; st.M = M;
    st.cs[0] = s * cx; st.cs[1] = s * cy; st.cs[2] = s * cz;
    st.co[0] = cx; st.co[1] = cy; st.co[2] = cz;
  }
  if (SE) {   // sides are (ref, fit) here
    const int s0 = pb.ref_srng[2 * pair], Ns = pb.ref_srng[2 * pair + 1];
    const int s1 = pb.fit_srng[2 * pair], Ms = pb.fit_srng[2 * pair + 1];
    const int a0 = pb.ref_arng[2 * pair], Na = pb.ref_arng[2 * pair + 1];
    const int a1 = pb.fit_arng[2 * pair], Ma = pb.fit_arng[2 * pair + 1];
    const float4* gsr = reinterpret_cast<const float4*>(pb.ref_srf) + s0;
    const float4* gsf = reinterpret_cast<const float4*>(pb.fit_srf) + s1;
    const float4* gar = reinterpret_cast<const float4*>(pb.ref_atm) + a0;
    const float4* gaf = reinterpret_cast<const float4*>(pb.fit_atm) + a1;
    float4* lsr = lds + lp.srf_r;
    float4* lsf = lds + lp.srf_f;
    float4* lar = lds + lp.atm_r;
    float4* laf = lds + lp.atm_f;
    float* lbr = reinterpret_cast<float*>(lds + lp.rb_r);
    float* lbf = reinterpret_cast<float*>(lds + lp.rb_f);
    const float cx = st.co[0], cy = st.co[1], cz = st.co[2], pr = ks.se_probe2;
    for (int i = tid; i < Ns; i += nt) {
      const float4 v = gsr[i];
      lsr[i] = make_float4(v.x - cx, v.y - cy, v.z - cz, v.w);
    }
    for (int i = tid; i < Ms; i += nt) lsf[i] = gsf[i];
    for (int i = tid; i < Na; i += nt) {
      const float4 v = gar[i];
      lar[i] = make_float4(v.x - cx, v.y - cy, v.z - cz, v.w);
      const float r = pb.ref_rad[a0 + i] + pr;
      lbr[i] = r * r;
    }
    for (int i = tid; i < Ma; i += nt) {
      laf[i] = gaf[i];
      const float r = pb.fit_rad[a1 + i] + pr;
      lbf[i] = r * r;
    }
    st.srf_r = lsr; st.srf_f = lsf; st.atm_r = lar; st.atm_f = laf; st.rb_r = lbr; st.rb_f = lbf;
    st.Ns = Ns; st.Ms = Ms; st.Na = Na; st.Ma = Ma;
  }
  if (PH) stage_pharm(pb, lp, pair, ref_side_a, fit_side_b, lds, st);
  return st;
}

// Similarity of one pose from its channel overlaps and, with grad, dS/dR and dS/dt.
template <typename T>
__device__ __forceinline__ T similarity(const KScore& ks, T Os, T Op, const double* so, bool has_sh,
                                        bool has_ph, T& ws, T& wp) {
  // ws, wp receive dS/dO of the shape and pharm channel (the ShaEP channel for ESP_COMBO).
  T S = 0;
  ws = 0; wp = 0;
  if (ks.mode == FSS_MODE_ESP_COMBO) {   // so[2] + so[3] = surface point counts
    const T w = sizeof(T) == 8 ? (T)ks.se_w_d : (T)ks.se_w;
    const T U = (T)so[0] + (T)so[1];
    const T D = U - Os;
    const T nrm = (T)so[2] + (T)so[3];
    ws = ((T)1 - w) * U / (D * D);
    wp = w / nrm;
    return ((T)1 - w) * (Os / D) + w * (Op / nrm);
  }
  const T wc = sizeof(T) == 8 ? (T)ks.w_d : (T)ks.w;
  const T sg = sizeof(T) == 8 ? (T)ks.sigma_d : (T)ks.sigma;
  const T chs = ks.mode == FSS_MODE_SHAPE_PHARM ? (T)1 - wc : (T)1;
  const T chp = ks.mode == FSS_MODE_SHAPE_PHARM ? wc : (T)1;
  if (has_sh) {
    const T U = (T)so[0] + (T)so[1];
    const T iD = (T)1 / (U - Os);
    S += chs * (Os * iD);
    ws = chs * U * (iD * iD);
  }
  if (has_ph) {
    if (ks.sim == FSS_SIM_TANIMOTO) {
      const T U = (T)so[2] + (T)so[3];
      const T iD = (T)1 / (U - Op);
      S += chp * (Op * iD);
      wp = chp * U * (iD * iD);
    } else {
      const T D = sg * (T)so[2] + ((T)1 - sg) * (T)so[3];
      const T r = Op / D;
      S += chp * (r < (T)1 ? r : (T)1);
      wp = Op < D ? chp / D : (T)0;
    }
  }
  return S;
}

// Per-pose loss gradient tail: dL/dR, dL/dt -> [dL/dq_raw, dL/dt] / B (projection in FP64).
// iq = 1 / |q_raw| as rotation_fp64 computed it.
__device__ __forceinline__ void se3_grad(const float gR[9], const float gT[3], const double qh[4],
                                         double iq, float invB, float g[7]) {
  double gq[4];
  project_to_quat(gR, qh, gq);
  const double dot = qh[0] * gq[0] + qh[1] * gq[1] + qh[2] * gq[2] + qh[3] * gq[3];
#pragma unroll
  for (int k = 0; k < 4; ++k) g[k] = (float)((gq[k] - qh[k] * dot) * iq) * invB;
#pragma unroll
  for (int k = 0; k < 3; ++k) g[4 + k] = gT[k] * invB;
}

// Sum an Acc over the workgroup (one value set per wave through LDS); result in thread 0.
__device__ __forceinline__ void block_sum(Acc& a, float* red, bool grad) {
  group_sum(a, 64, grad);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (l == 0) {
    red[w * 13] = a.V;
    for (int i = 0; i < 3; ++i) red[w * 13 + 1 + i] = a.gt[i];
    for (int i = 0; i < 9; ++i) red[w * 13 + 4 + i] = a.G[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Acc s; s.zero();
    for (int k = 0; k < nw; ++k) {
      s.V += red[k * 13];
      for (int i = 0; i < 3; ++i) s.gt[i] += red[k * 13 + 1 + i];
      for (int i = 0; i < 9; ++i) s.G[i] += red[k * 13 + 4 + i];
    }
    a = s;
  }
}

__device__ __forceinline__ double block_sum_d(double v, double* red) {
  v = group_sum(v, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < nw; ++k) s += red[k];
  return s;
}

// ------------------------------------------------------------------------------------------------
// Self-overlaps, FP64, one workgroup per (pair, side).
// ------------------------------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(kPoseBlock) void k_self_overlap(FssPairBatch pb, KScore ks, double* out) {
  constexpr bool SH = MODE != FSS_MODE_PHARM;
  constexpr bool ESP = MODE == FSS_MODE_SHAPE_ESP;
  constexpr bool PH = MODE == FSS_MODE_PHARM || MODE == FSS_MODE_SHAPE_PHARM;
  __shared__ double red[kPoseBlock / 64];
  const int pair = blockIdx.x, side = blockIdx.y;
  const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double z[3] = {0, 0, 0};
  double vs = 0.0, vp = 0.0;
  if (SH) {
    const int* rg = side ? pb.fit_rng : pb.ref_rng;
    const float4* p = reinterpret_cast<const float4*>(side ? pb.fit_pts : pb.ref_pts) + rg[2 * pair];
    const int n = rg[2 * pair + 1];
    vs = shape_value_f64<ESP>(p, n, p, n, threadIdx.x, blockDim.x, I, z, 0.5 * ks.alpha_d, 1.0 / ks.lam_d);
    vs = block_sum_d(vs, red) * pow(M_PI / (2.0 * ks.alpha_d), 1.5);
  }
  if (PH) {
    const int* rg = side ? pb.fit_prng : pb.ref_prng;
    const int o = rg[2 * pair], n = rg[2 * pair + 1];
    const float4* a = reinterpret_cast<const float4*>(side ? pb.fit_anc : pb.ref_anc) + o;
    const float4* v = reinterpret_cast<const float4*>(side ? pb.fit_vec : pb.ref_vec) + o;
    const int32_t* ty = (side ? pb.fit_typ : pb.ref_typ) + o;
    vp = pharm_value_f64(a, v, ty, n, a, v, ty, n, threadIdx.x, blockDim.x, I, z, ks.ph_alpha, ks.ph_k, ks.ph_cat);
    vp = block_sum_d(vp, red);
  }
  if (threadIdx.x == 0) {
    const int* srng = side ? pb.fit_srng : pb.ref_srng;
    out[4 * pair + side] = SH ? vs : 0.0;
    out[4 * pair + 2 + side] = PH ? vp : (MODE == FSS_MODE_ESP_COMBO ? (double)srng[2 * pair + 1] : 0.0);
  }
}

// ------------------------------------------------------------------------------------------------
// One workgroup per pose.  OUT == 0: channel overlaps and gradients; OUT == 1: loss + SE(3) grad.
// ------------------------------------------------------------------------------------------------
template <int MODE, bool GRAD, int OUT>
__global__ __launch_bounds__(kPoseBlock) void k_pose_grad(FssPairBatch pb, KScore ks, LdsPlan lp,
                                                          const int32_t* __restrict__ pose_pair,
                                                          const float* __restrict__ qin,
                                                          const float* __restrict__ tin,
                                                          const double* __restrict__ self_ov,
                                                          float invB, float* V, float* gR,
                                                          float* gT, float* gQ, float* loss,
                                                          float* grad) {
  constexpr bool SH = MODE != FSS_MODE_PHARM;
  constexpr bool ESP = MODE == FSS_MODE_SHAPE_ESP;
  constexpr bool PH = MODE == FSS_MODE_PHARM || MODE == FSS_MODE_SHAPE_PHARM;
  extern __shared__ float4 lds[];
  constexpr bool SE = MODE == FSS_MODE_ESP_COMBO;
  __shared__ float red[2][(kPoseBlock / 64) * 13];
  const int pose = blockIdx.x;
  const int pair = pose_pair[pose];
  // streamed shape clouds when the plan says so (tileable modes whose clouds exceed LDS)
  bool tl = false;
  if constexpr (tileable(MODE)) tl = lp.tile_ref > 0;
  TileState tiles;
  const Staged st = tl ? stage_tiled<ESP, PH>(pb, ks, lp, pair, lds)
                       : stage_pair<SH, ESP, PH, true, false, SE>(pb, ks, lp, pair, 0, 1, lds);
  float q[4] = {qin[4 * pose], qin[4 * pose + 1], qin[4 * pose + 2], qin[4 * pose + 3]};
  const float t[3] = {tin[3 * pose], tin[3 * pose + 1], tin[3 * pose + 2]};
  double qh[4], qn, iq, Rd[9];
  rotation_fp64(q, qh, qn, Rd, &iq);
  float R[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) R[i] = (float)Rd[i];
  __syncthreads();
  Acc sh, ph;
  sh.zero();
  ph.zero();
  if (SH) {
    const float ts[3] = {fmaf(t[0], ks.s, -st.cs[0]), fmaf(t[1], ks.s, -st.cs[1]), fmaf(t[2], ks.s, -st.cs[2])};
    if (tl) shape_accum_tiled<ESP, GRAD>(st, ks, lp, tiles, true, threadIdx.x, blockDim.x, R, ts, sh);
    else shape_accum<ESP, GRAD>(st.ref, st.refc, st.N, st.fit, st.M, threadIdx.x, blockDim.x, R, ts, sh);
    block_sum(sh, red[0], GRAD);
  }
  if (PH) {
    pharm_accum<GRAD>(st.ranc, st.rvec, st.poff, st.fanc, st.fvec, st.Mp, threadIdx.x, blockDim.x, R, t, ks.tab, ph);
    block_sum(ph, red[1], GRAD);
  }
  if (SE) {   // the ShaEP comparison takes the second channel slot
    const float tr[3] = {t[0] - st.co[0], t[1] - st.co[1], t[2] - st.co[2]};
    shaep_accum<GRAD>(st.srf_r, st.Ns, st.atm_f, st.rb_f, st.Ma, st.srf_f, st.Ms, st.atm_r, st.rb_r, st.Na,
                      threadIdx.x, blockDim.x, R, tr, ks.se, ph);
    block_sum(ph, red[1], GRAD);
  }
  if (threadIdx.x != 0) return;
  if (SH) shape_finish(sh, ks.K, ks.alpha, ks.s, GRAD);
  if (OUT == 0) {
    const Acc* ch[2] = {&sh, &ph};
    for (int c = 0; c < 2; ++c) {
      V[2 * pose + c] = ch[c]->V;
      if (GRAD) {
        for (int i = 0; i < 9; ++i) gR[18 * pose + 9 * c + i] = ch[c]->G[i];
        for (int i = 0; i < 3; ++i) gT[6 * pose + 3 * c + i] = ch[c]->gt[i];
        double g4[4];
        project_to_quat(ch[c]->G, qh, g4);
        for (int i = 0; i < 4; ++i) gQ[8 * pose + 4 * c + i] = (float)g4[i];
      }
    }
  } else {
    float ws, wp;
    const float S = similarity<float>(ks, sh.V, ph.V, self_ov + 4 * pair, SH, PH, ws, wp);
    float lR[9], lT[3];
    for (int i = 0; i < 9; ++i) lR[i] = -(ws * sh.G[i] + wp * ph.G[i]);
    for (int i = 0; i < 3; ++i) lT[i] = -(ws * sh.gt[i] + wp * ph.gt[i]);
    float g[7];
    se3_grad(lR, lT, qh, iq, invB, g);
    loss[pose] = 1.f - S;
    for (int i = 0; i < 7; ++i) grad[7 * pose + i] = g[i];
  }
}

// ------------------------------------------------------------------------------------------------
// torch.optim.Adam step (single-tensor form: lerp for m, bias corrections in double).
// ------------------------------------------------------------------------------------------------
// Parameter update of torch.optim.Adam (bias-corrected, eps outside the square root) with the
// step's 1 / sqrt(1 - beta2^step) precomputed and the hardware square root and reciprocal (about
// 1 ulp each) in place of IEEE division: a few instructions per parameter instead of two
// division sequences.  Shared by the persistent and the stepped optimiser.
__device__ __forceinline__ float adam_update(float p, float m, float v, float step_size, float ibc2s, float eps) {
  const float den = fmaf(__builtin_amdgcn_sqrtf(v), ibc2s, eps);
  return p - step_size * (m * __builtin_amdgcn_rcpf(den));
}

__global__ void k_adam(int n, float* q, float* t, const float* __restrict__ g, float* m, float* v,
                       int step, float lr, float b1, float b2, float eps,
                       const int32_t* __restrict__ active, int B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * 7) return;
  const int row = i / 7, c = i % 7;
  if (active && !active[row / B]) return;
  float* p = c < 4 ? q + 4 * row + c : t + 3 * row + (c - 4);
  const double bc1 = 1.0 - pow((double)b1, (double)step);
  const double bc2 = 1.0 - pow((double)b2, (double)step);
  const float step_size = (float)((double)lr / bc1);
  const float ibc2s = (float)(1.0 / sqrt(bc2));
  const float gi = g[i];
  float mi = m[i], vi = v[i];
  mi = fmaf(1.f - b1, gi - mi, mi);
  vi = fmaf(1.f - b2, gi * gi, vi * b2);
  m[i] = mi;
  v[i] = vi;
  *p = adam_update(*p, mi, vi, step_size, ibc2s, eps);
}

// The reference's accelerated-driver Adam on unit quaternions: one thread per pose row, moments
// without bias correction, eps inside the square root, optional tangent projection of the
// quaternion gradient (dq - q (dq . q)), then renormalisation of q.
__global__ void k_adam_qt(int n, float* q, float* t, const float* __restrict__ dQ,
                          const float* __restrict__ dT, float* mq, float* vq, float* mt, float* vt,
                          float lr, float b1, float c1, float b2, float c2, float eps, int proj) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float qq[4], g[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { qq[k] = q[4 * i + k]; g[k] = dQ[4 * i + k]; }
  if (proj) {
    const float radial = fmaf(g[0], qq[0], fmaf(g[1], qq[1], fmaf(g[2], qq[2], g[3] * qq[3])));
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = fmaf(-qq[k], radial, g[k]);
  }
  float nrm2 = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float m = fmaf(b1, mq[4 * i + k], c1 * g[k]);
    const float v = fmaf(b2, vq[4 * i + k], c2 * g[k] * g[k]);
    mq[4 * i + k] = m;
    vq[4 * i + k] = v;
    qq[k] -= lr * m / sqrtf(v + eps);
    nrm2 = fmaf(qq[k], qq[k], nrm2);
  }
  const float nrm = sqrtf(nrm2);
#pragma unroll
  for (int k = 0; k < 4; ++k) q[4 * i + k] = qq[k] / nrm;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float gk = dT[3 * i + k];
    const float m = fmaf(b1, mt[3 * i + k], c1 * gk);
    const float v = fmaf(b2, vt[3 * i + k], c2 * gk * gk);
    mt[3 * i + k] = m;
    vt[3 * i + k] = v;
    t[3 * i + k] -= lr * m / sqrtf(v + eps);
  }
}

// ------------------------------------------------------------------------------------------------
// Persistent optimiser.
// ------------------------------------------------------------------------------------------------
// Adam's bias-corrected step factors of the first kSchedSteps steps, formed on the host in FP64
// exactly as the kernel would (the same correctly rounded products, quotient and square root),
// so a step reads two floats from the kernel arguments instead of an FP64 divide and square root.
constexpr int kSchedSteps = 256;

struct AlignK {
  int B, max_steps;
  float lr, b1, b2, eps;
  float sched[2 * kSchedSteps];   // step s (1-based): lr / (1 - b1^s), 1 / sqrt(1 - b2^s)
  double b1p_tail, b2p_tail;      // b1^kSchedSteps, b2^kSchedSteps: the products the kernel continues
  float es_tol;
  int es_pat;
  int lpp, rpw, wg_per_pair;
  int groups;                 // lane groups per workgroup; each runs ceil(rpw / groups) restarts
  int red_lds;                // lane groups of a non-power-of-two width sum through LDS
  int xwave;                  // lane groups of several wavefronts (streamed path) sum through LDS
  int ppw;                    // pairs per workgroup (> 1: each with all its restarts)
  int region;                 // float4s between the LDS regions of a workgroup's pairs
  const int32_t* order;
  const double* self_ov;
  float* q;
  float* t;
  double* pose_score;
  int32_t* n_steps;
};

// Every instance is held to 128 VGPRs (4 wavefronts per SIMD).  The MFMA loop is latency-bound,
// and shape + pharm with early stop would otherwise take 133 VGPRs: with 4 spilled dwords it
// runs 6 % faster at 4 wavefronts than at 3 (16384 pairs, 38.2 -> 35.8 ms).  The several-pairs
// instances spill 68-168 B at 128 VGPRs, outside the pair-term loops; at 3 wavefronts they spill
// nothing and screen 3 % slower (vol, 10 restarts: 4.17M -> 4.05M pairs/s).
// MP: several pairs per workgroup (AlignK::ppw > 1), a separate instance so the one-pair
// instances keep their pair's staging offsets and counts uniform (scalar) across the workgroup.
template <int MODE, bool ES, int MAXT, int VAR, bool MP = false>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(4)))
void k_align(FssPairBatch pb, KScore ks, LdsPlan lp, AlignK ak) {
  constexpr bool EXPD = VAR != 1;
  constexpr bool MF = VAR == 3;
  constexpr bool TL = VAR == 4;   // streamed shape clouds (tileable modes, one pair per workgroup)
  constexpr bool SE = MODE == FSS_MODE_ESP_COMBO;
  constexpr bool SH = MODE != FSS_MODE_PHARM;
  constexpr bool ESP = MODE == FSS_MODE_SHAPE_ESP;
  constexpr bool PH = MODE == FSS_MODE_PHARM || MODE == FSS_MODE_SHAPE_PHARM;
  static_assert(!TL || (tileable(MODE) && !MP), "the streamed path runs the tileable modes, one pair per workgroup");
  extern __shared__ float4 lds[];
  const int lpp = ak.lpp, groups = ak.groups, ppw = MP ? ak.ppw : 1;
  // ppw pairs per workgroup (each with every restart in the workgroup): lane group g works on
  // the workgroup's pair g / groups; with one pair the workgroup may hold a block of restarts.
  const int grp_all = threadIdx.x / lpp, j = threadIdx.x - grp_all * lpp;
  const int pl = MP ? grp_all / groups : (grp_all < groups ? 0 : 1);
  const int grp = MP ? grp_all - pl * groups : grp_all;
  const int slot0 = ppw > 1 ? blockIdx.x * ppw : blockIdx.x / ak.wg_per_pair;
  const int blk = ppw > 1 ? 0 : blockIdx.x - slot0 * ak.wg_per_pair;
  // lanes past the workgroup's lane groups (and groups of a missing last pair) have no pair
  const bool has_pair = pl < ppw && (ppw == 1 || slot0 + pl < pb.n_pairs);
  const int pl_buf = ppw == 1 ? 0 : pl;     // one pair: every lane reads that pair's buffers
  // one pair: the workgroup's pair for every lane (uniform); several: the lane's own
  const int pair = !MP ? (ak.order ? ak.order[slot0] : slot0)
                       : has_pair ? (ak.order ? ak.order[slot0 + pl] : slot0 + pl) : 0;
  Staged st{};
  TileState tiles;
  if (TL) {
    st = stage_tiled<ESP, PH>(pb, ks, lp, pair, lds);
  } else if (ppw == 1) {
    st = stage_pair<SH, ESP, PH, EXPD, MF, SE>(pb, ks, lp, pair, 0, 1, lds);
  } else {
    for (int k = 0; k < ppw && slot0 + k < pb.n_pairs; ++k) {   // the same trip count everywhere
      const int pk = ak.order ? ak.order[slot0 + k] : slot0 + k;
      const Staged sk = stage_pair<SH, ESP, PH, EXPD, MF, SE>(pb, ks, lp, pk, 0, 1, lds + k * ak.region);
      if (k == pl) st = sk;
    }
  }
  // per pair of the workgroup: 2 * rpw doubles of restart losses (ES only), then rpw slots
  float4* const tail = lds + (MP ? ppw * ak.region : lp.end);
  double* lossbuf = reinterpret_cast<double*>(tail) + (ES ? pl_buf * 2 * ak.rpw : 0);
  // Each lane group runs the restarts grp, grp + groups, ... of the workgroup's block, one after
  // the other in every step.  A restart's Adam moments (and, when its group runs more than one,
  // its parameters) live in an LDS slot: m at 0..6, v at 7..13, and q and t at 14..20 in the
  // 24-float slots of groups running several restarts; lane 0 of the group writes, the group's
  // lanes read (within one wavefront the DS operations complete in order; a group of several
  // wavefronts, streamed path only, reads and writes on either side of a barrier).  With one
  // restart per group the parameters stay in registers across steps and a slot is 14 floats.
  const int rpg = (ak.rpw + groups - 1) / groups;
  const double* so = ak.self_ov + 4 * pair;
  const float invB = 1.0f / (float)ak.B;
  float* slot_base = reinterpret_cast<float*>(reinterpret_cast<double*>(tail) + (ES ? ppw * 2 * ak.rpw : 0));
  // a restart's slot: Adam moments, and its parameters when its group runs several restarts
  const int sls = rpg > 1 ? 24 : 14;
  float* slots = slot_base + pl_buf * sls * ak.rpw;
  // lane groups of a width other than a power of two sum through LDS when the plan reserved it
  float* red = ak.red_lds || ak.xwave ? slot_base + ppw * sls * ak.rpw : nullptr;
  const int nl = ppw * groups * lpp;
  auto local = [&](int i) { return grp + i * groups; };
  auto runs = [&](int rl) { return has_pair && rl < ak.rpw && blk * ak.rpw + rl < ak.B; };
  auto pose_of = [&](int rl) { return (long)pair * ak.B + blk * ak.rpw + rl; };
  float p[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) p[k] = 0.f;
  for (int i = 0; i < rpg; ++i) {
    const int rl = local(i);
    if (!runs(rl)) continue;
    const long pose = pose_of(rl);
    float* sl = slots + sls * rl;
    if (rpg == 1) {
#pragma unroll
      for (int k = 0; k < 4; ++k) p[k] = ak.q[4 * pose + k];
#pragma unroll
      for (int k = 0; k < 3; ++k) p[4 + k] = ak.t[3 * pose + k];
    }
    if (j == 0) {
      for (int k = 0; k < 14; ++k) sl[k] = 0.f;
      if (rpg > 1) {
        for (int k = 0; k < 4; ++k) sl[14 + k] = ak.q[4 * pose + k];
        for (int k = 0; k < 3; ++k) sl[18 + k] = ak.t[3 * pose + k];
      }
    }
  }
  __syncthreads();

  // early-stop reference loss before the first step: 1.0 for the analytic optimisers
  // (_torch_analytical.py:110, 283, 420), +inf for the ShaEP one (alignment/_torch.py:922)
  double b1p = ak.b1p_tail, b2p = ak.b2p_tail, last = SE ? (double)INFINITY : 1.0;
  int counter = 0, done = ak.max_steps;
  // this lane's pair still steps: with one pair every lane (grouped or not) decides alike
  bool active = ppw == 1 || has_pair;
  for (int step = 1; step <= ak.max_steps; ++step) {
    float step_size, ibc2s;
    if (step <= kSchedSteps) {
      step_size = ak.sched[2 * step - 2];
      ibc2s = ak.sched[2 * step - 1];
    } else {
      b1p *= (double)ak.b1;
      b2p *= (double)ak.b2;
      step_size = (float)((double)ak.lr / (1.0 - b1p));
      ibc2s = (float)(1.0 / sqrt(1.0 - b2p));
    }
    double* buf = lossbuf + (step & 1) * ak.rpw;
    for (int i = 0; i < rpg; ++i) {
      const int rl = local(i);
      // the streamed path's lanes without a restart still stage tiles and meet their barriers
      const bool work = active && runs(rl);
      if (!TL && !work) continue;
      float* sl = slots + sls * (work ? rl : 0);
      if (rpg > 1 && work) {
#pragma unroll
        for (int k = 0; k < 7; ++k) p[k] = sl[14 + k];
      }
      double qh[4], qn, iq, Rd[9];
      rotation_fp64(p, qh, qn, Rd, &iq);
      float R[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) R[k] = (float)Rd[k];
      Acc sh, ph;
      sh.zero();
      ph.zero();
      if (SH) {
        const float ts[3] = {fmaf(p[4], ks.s, -st.cs[0]), fmaf(p[5], ks.s, -st.cs[1]), fmaf(p[6], ks.s, -st.cs[2])};
        if (TL) shape_accum_tiled<ESP, true>(st, ks, lp, tiles, work, j, lpp, R, ts, sh);
        else if (MF) shape_accum_mfma<ESP>(st.ref, st.xq, st.T, st.fit, st.M, R, ts, sh);
        else if (EXPD) shape_accum<ESP, true>(st.ref, st.refc, st.N, st.fit, st.M, j, lpp, R, ts, sh);
        else shape_accum_direct<ESP, true>(st.ref, st.N, st.fit, st.M, j, lpp, R, ts, sh);
        // lane groups of several wavefronts: every lane goes on to the slot barrier below
        if (TL && ak.xwave) pose_sum_waves(sh, lpp, red);
        if (TL && !ak.xwave && !work) continue;
        if (!TL || !ak.xwave) pose_sum(sh, lpp, j, true, grp_all, nl, red);
        shape_finish(sh, ks.K, ks.alpha, ks.s, true);
      }
      if (PH) {
        // streamed path, lane groups of several wavefronts: every lane meets the sums' barriers
        if (!TL || work) pharm_accum<true>(st.ranc, st.rvec, st.poff, st.fanc, st.fvec, st.Mp, j, lpp, R, p + 4, ks.tab, ph);
        if (TL && ak.xwave) pose_sum_waves(ph, lpp, red);
        else pose_sum(ph, lpp, j, true, grp_all, nl, red);
      }
      if (SE) {
        const float tr[3] = {p[4] - st.co[0], p[5] - st.co[1], p[6] - st.co[2]};
        shaep_accum<true>(st.srf_r, st.Ns, st.atm_f, st.rb_f, st.Ma, st.srf_f, st.Ms, st.atm_r, st.rb_r, st.Na,
                          j, lpp, R, tr, ks.se, ph);
        pose_sum(ph, lpp, j, true, grp_all, nl, red);
      }
      float ws, wp;
      const float S = similarity<float>(ks, sh.V, ph.V, so, SH, PH, ws, wp);
      const float loss = 1.f - S;
      float lR[9], lT[3];
#pragma unroll
      for (int k = 0; k < 9; ++k) lR[k] = -(ws * sh.G[k] + wp * ph.G[k]);
#pragma unroll
      for (int k = 0; k < 3; ++k) lT[k] = -(ws * sh.gt[k] + wp * ph.gt[k]);
      float g[7];
      se3_grad(lR, lT, qh, iq, invB, g);
      if constexpr (TL) {
        // A lane group of several wavefronts reads its restart's slot in every wavefront: all of
        // them read before the group's lane 0 rewrites it.
        float mk[7], vk[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
          mk[k] = fmaf(1.f - ak.b1, g[k] - sl[k], sl[k]);
          vk[k] = fmaf(1.f - ak.b2, g[k] * g[k], sl[7 + k] * ak.b2);
          p[k] = adam_update(p[k], mk[k], vk[k], step_size, ibc2s, ak.eps);
        }
        if (ak.xwave) __syncthreads();
        if (!work) continue;
        if (j == 0) {
#pragma unroll
          for (int k = 0; k < 7; ++k) {
            sl[k] = mk[k];
            sl[7 + k] = vk[k];
            if (rpg > 1) sl[14 + k] = p[k];
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < 7; ++k) {
          const float mk = fmaf(1.f - ak.b1, g[k] - sl[k], sl[k]);
          const float vk = fmaf(1.f - ak.b2, g[k] * g[k], sl[7 + k] * ak.b2);
          p[k] = adam_update(p[k], mk, vk, step_size, ibc2s, ak.eps);
          if (j == 0) {
            sl[k] = mk;
            sl[7 + k] = vk;
            if (rpg > 1) sl[14 + k] = p[k];
          }
        }
      }
      if (ES && j == 0) buf[rl] = (double)loss;
    }
    if (ES) {
      __syncthreads();
      if (active) {
        double L = 0.0;
        for (int i = 0; i < ak.B; ++i) L += buf[i];
        L /= (double)ak.B;
        if (fabs(L - last) > (double)ak.es_tol) counter = 0;
        else ++counter;
        last = L;
        if (counter > ak.es_pat) { done = step; active = false; }
      }
      // one pair: every lane decided alike; several: step on while any pair of the workgroup runs
      if (ppw == 1 ? !active : !__syncthreads_or(active)) break;
    }
  }

  for (int i = 0; i < rpg; ++i) {
    const int rl = local(i);
    if (!runs(rl) || j != 0) continue;
    const long pose = pose_of(rl);
    const float* src = rpg > 1 ? slots + sls * rl + 14 : p;
#pragma unroll
    for (int k = 0; k < 4; ++k) ak.q[4 * pose + k] = src[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) ak.t[3 * pose + k] = src[4 + k];
  }
  if (has_pair && grp == 0 && j == 0 && blk == 0) ak.n_steps[pair] = done;
}

// ------------------------------------------------------------------------------------------------
// FP64 similarity of every pose, one wave per pose (4 poses per workgroup), straight from the
// packed unscaled arrays.  pose_pair == NULL: pair = pose / B.
// ------------------------------------------------------------------------------------------------
// thr != NULL: only poses whose current score (the FP32 pre-pass) is >= thr[pair] are re-scored.
template <int MODE>
__global__ __launch_bounds__(256) void k_pose_score(FssPairBatch pb, KScore ks, int n_poses, int B,
                                                    const int32_t* __restrict__ pose_pair,
                                                    const double* __restrict__ self_ov,
                                                    const float* __restrict__ qin,
                                                    const float* __restrict__ tin, double* score,
                                                    const double* __restrict__ thr) {
  constexpr bool SH = MODE != FSS_MODE_PHARM;
  constexpr bool ESP = MODE == FSS_MODE_SHAPE_ESP;
  constexpr bool PH = MODE == FSS_MODE_PHARM || MODE == FSS_MODE_SHAPE_PHARM;
  // one pose per wavefront: wave-uniform, so the clouds are read with scalar loads
  const int pose = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int j = threadIdx.x & 63;
  if (pose >= n_poses) return;
  const int pair = __builtin_amdgcn_readfirstlane(pose_pair ? pose_pair[pose] : pose / B);
  if (thr && !(score[pose] >= thr[pair])) return;
  const float q[4] = {qin[4 * (long)pose], qin[4 * (long)pose + 1], qin[4 * (long)pose + 2], qin[4 * (long)pose + 3]};
  double qh[4], qn, Rd[9];
  rotation_fp64(q, qh, qn, Rd);
  const double td[3] = {tin[3 * (long)pose], tin[3 * (long)pose + 1], tin[3 * (long)pose + 2]};
  double vs = 0.0, vp = 0.0;
  if (SH) {
    const float4* gref = reinterpret_cast<const float4*>(pb.ref_pts) + pb.ref_rng[2 * pair];
    const float4* gfit = reinterpret_cast<const float4*>(pb.fit_pts) + pb.fit_rng[2 * pair];
    vs = shape_value_f64<ESP>(gref, pb.ref_rng[2 * pair + 1], gfit, pb.fit_rng[2 * pair + 1], j, 64, Rd, td,
                              0.5 * ks.alpha_d, 1.0 / ks.lam_d);
    vs = group_sum(vs, 64) * pow(M_PI / (2.0 * ks.alpha_d), 1.5);
  }
  if (PH) {
    const int ro = pb.ref_prng[2 * pair], fo = pb.fit_prng[2 * pair];
    vp = pharm_value_f64(reinterpret_cast<const float4*>(pb.ref_anc) + ro,
                         reinterpret_cast<const float4*>(pb.ref_vec) + ro, pb.ref_typ + ro, pb.ref_prng[2 * pair + 1],
                         reinterpret_cast<const float4*>(pb.fit_anc) + fo,
                         reinterpret_cast<const float4*>(pb.fit_vec) + fo, pb.fit_typ + fo, pb.fit_prng[2 * pair + 1],
                         j, 64, Rd, td, ks.ph_alpha, ks.ph_k, ks.ph_cat);
    vp = group_sum(vp, 64);
  }
  if (MODE == FSS_MODE_ESP_COMBO) {
    const int s0 = pb.ref_srng[2 * pair], s1 = pb.fit_srng[2 * pair];
    const int a0 = pb.ref_arng[2 * pair], a1 = pb.fit_arng[2 * pair];
    vp = shaep_value_f64(reinterpret_cast<const float4*>(pb.ref_srf) + s0, pb.ref_srng[2 * pair + 1],
                         reinterpret_cast<const float4*>(pb.fit_atm) + a1, pb.fit_rad + a1, pb.fit_arng[2 * pair + 1],
                         reinterpret_cast<const float4*>(pb.fit_srf) + s1, pb.fit_srng[2 * pair + 1],
                         reinterpret_cast<const float4*>(pb.ref_atm) + a0, pb.ref_rad + a0, pb.ref_arng[2 * pair + 1],
                         j, 64, Rd, td, ks.se_coul_d, ks.se_lam_d, ks.se_probe_d);
    vp = group_sum(vp, 64);
  }
  double ws, wp;
  const double S = similarity<double>(ks, vs, vp, self_ov + 4 * pair, SH, PH, ws, wp);
  if (j == 0) score[pose] = S;
}

// Re-scoring pre-pass: the FP32 similarity of every final pose (one wave per pose, direct form),
// then per pair the threshold max - kRescoreMargin over its restarts.  Only restarts at or above
// it get the FP64 score; the others keep the FP32 one, whose error (about 1e-6) is far below the
// margin, so the FP64 argmax over the restarts is unchanged.  Every mode but ShaEP.
constexpr double kRescoreMargin = 1e-3;

template <int MODE>
__global__ __launch_bounds__(256) void k_pose_score_f32(FssPairBatch pb, KScore ks, int n_poses, int B,
                                                        const double* __restrict__ self_ov,
                                                        const float* __restrict__ qin,
                                                        const float* __restrict__ tin, double* score) {
  static_assert(MODE != FSS_MODE_ESP_COMBO, "the ShaEP mode scores every pose in FP64");
  constexpr bool SH = MODE != FSS_MODE_PHARM;
  constexpr bool ESP = MODE == FSS_MODE_SHAPE_ESP;
  constexpr bool PH = MODE == FSS_MODE_PHARM || MODE == FSS_MODE_SHAPE_PHARM;
  const int pose = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int j = threadIdx.x & 63;
  if (pose >= n_poses) return;
  const int pair = pose / B;
  const float q[4] = {qin[4 * (long)pose], qin[4 * (long)pose + 1], qin[4 * (long)pose + 2], qin[4 * (long)pose + 3]};
  double qh[4], qn, Rd[9];
  rotation_fp64(q, qh, qn, Rd);
  float R[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) R[i] = (float)Rd[i];
  const float t[3] = {tin[3 * (long)pose], tin[3 * (long)pose + 1], tin[3 * (long)pose + 2]};
  float v = 0.f, vp = 0.f;
  if (SH) {
    const float4* gref = reinterpret_cast<const float4*>(pb.ref_pts) + pb.ref_rng[2 * pair];
    const float4* gfit = reinterpret_cast<const float4*>(pb.fit_pts) + pb.fit_rng[2 * pair];
    v = shape_value_f32<ESP>(gref, pb.ref_rng[2 * pair + 1], gfit, pb.fit_rng[2 * pair + 1], j, 64, R, t,
                             -ks.s * ks.s, ESP ? -ks.sc * ks.sc : 0.f);
    v = group_sum(v, 64) * ks.K;
  }
  if (PH) {
    const int ro = pb.ref_prng[2 * pair], fo = pb.fit_prng[2 * pair];
    vp = pharm_value_f32(reinterpret_cast<const float4*>(pb.ref_anc) + ro,
                         reinterpret_cast<const float4*>(pb.ref_vec) + ro, pb.ref_typ + ro, pb.ref_prng[2 * pair + 1],
                         reinterpret_cast<const float4*>(pb.fit_anc) + fo,
                         reinterpret_cast<const float4*>(pb.fit_vec) + fo, pb.fit_typ + fo, pb.fit_prng[2 * pair + 1],
                         j, 64, R, t, ks.tab);
    vp = group_sum(vp, 64);
  }
  float ws, wp;
  const float S = similarity<float>(ks, v, vp, self_ov + 4 * pair, SH, PH, ws, wp);
  if (j == 0) score[pose] = (double)S;
}

// thr[pair] = (max over the pair's restarts of score) - kRescoreMargin, ignoring NaN scores.  A
// NaN FP32 score (NaN inputs: the FP64 score is NaN too) is never a candidate.
__global__ void k_rescore_threshold(int n_pairs, int B, const double* __restrict__ score, double* thr) {
  const int pair = blockIdx.x * blockDim.x + threadIdx.x;
  if (pair >= n_pairs) return;
  double m = -INFINITY;
  for (int r = 0; r < B; ++r) {
    const double v = score[(long)pair * B + r];
    if (v > m) m = v;
  }
  thr[pair] = m - kRescoreMargin;
}

// Per-pair early-stop bookkeeping of the stepped optimiser: mean loss over the pair's B poses,
// the |dL| <= tol counter, and deactivation (recording the step count) once it exceeds the
// patience.  State: last (n_pairs) double, counter / active / n_steps (n_pairs) int32.
__global__ void k_es_update(int n_pairs, int B, const float* __restrict__ loss, double* last,
                            int32_t* counter, int32_t* active, int32_t* n_steps, int step, double tol,
                            int patience) {
  const int pair = blockIdx.x * blockDim.x + threadIdx.x;
  if (pair >= n_pairs || !active[pair]) return;
  double L = 0.0;
  for (int r = 0; r < B; ++r) L += (double)loss[(long)pair * B + r];
  L /= (double)B;
  int c = counter[pair];
  c = fabs(L - last[pair]) > tol ? 0 : c + 1;
  counter[pair] = c;
  last[pair] = L;
  n_steps[pair] = step;
  if (c > patience) active[pair] = 0;
}

__global__ void k_align_finish(int n_pairs, int B, const float* __restrict__ q,
                               const float* __restrict__ t, const double* __restrict__ ps,
                               double* best_score, int32_t* best_idx, float* best_tf) {
  const int pair = blockIdx.x * blockDim.x + threadIdx.x;
  if (pair >= n_pairs) return;
  int bi = 0;
  double bs = ps[(long)pair * B];
  for (int r = 1; r < B; ++r) {
    const double s = ps[(long)pair * B + r];
    if (s > bs || (isnan(s) && !isnan(bs))) { bs = s; bi = r; }
  }
  const long pose = (long)pair * B + bi;
  const float qq[4] = {q[4 * pose], q[4 * pose + 1], q[4 * pose + 2], q[4 * pose + 3]};
  double qh[4], qn, Rd[9];
  rotation_fp64(qq, qh, qn, Rd);
  float* T = best_tf + 16 * (long)pair;
  for (int i = 0; i < 3; ++i) {
    for (int k = 0; k < 3; ++k) T[4 * i + k] = (float)Rd[3 * i + k];
    T[4 * i + 3] = t[3 * pose + i];
  }
  T[12] = 0.f; T[13] = 0.f; T[14] = 0.f; T[15] = 1.f;
  best_score[pair] = bs;
  best_idx[pair] = bi;
}

// ------------------------------------------------------------------------------------------------
// Differentiable similarity of fixed clouds (shape or ESP shape), FP32 or FP64.
//
// Both clouds of a pair are staged in LDS as 4-D points p = (s x, s y, s z, s_w w), with
// s = sqrt(c alpha / 2), s_w = sqrt(c / lam) (0 without ESP), c = log2(e) for FP32 (exp2) and 1
// for FP64 (exp), after subtracting the ref cloud's centroid (every sum is translation
// invariant).  Then w_ab E_ab = base^(-|p_a - p_b|^2): the ESP weight exp(-(w_a - w_b)^2 / lam)
// is the fourth coordinate of one 4-D Gaussian, and each term costs one exponent (expanded form
// in FP32, direct form in FP64), one exponential and the sums.
// ------------------------------------------------------------------------------------------------
template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

template <typename T> __device__ __forceinline__ T neg_exp(T r2);
template <> __device__ __forceinline__ float neg_exp<float>(float r2) { return fast_exp2(-r2); }
template <> __device__ __forceinline__ double neg_exp<double>(double r2) { return exp(-r2); }
template <typename T> constexpr T score_base() { return sizeof(T) == 4 ? (T)kLog2eD : (T)1; }

constexpr int kScoreBlock = 512;                 // threads of the score-as-loss kernels (8 wavefronts)
static_assert((kScoreBlock / 64 & (kScoreBlock / 64 - 1)) == 0, "the forward deals tiles by unit & (wavefronts - 1)");
constexpr int kScoreSlices = 4;                  // backward: lanes sharing one row
constexpr int kScoreRows = 64 / kScoreSlices;    // backward: rows per wavefront unit
static_assert(kScoreSlices == 4, "k_score_bwd combines the slices with two shuffles");

// FP32 instances take the exponent in the expanded form -|p - q|^2 = W_p + W_q + 2 p.q, with
// W = -|p|^2 staged beside each point: 1 add and 3-4 FMA per term in place of 4 differences and
// a squared norm.  The frame's origin is the ref cloud's centroid, so |p|^2 stays at molecule
// scale (a few tens) and the rounding of the cancelling terms is a few 1e-6 of a term.  FP64
// instances (gradcheck) take the direct form.
template <typename T> constexpr bool score_expanded() { return sizeof(T) == 4; }

// Origin of the scaled frame: the ref cloud's centroid (the fit's when the ref is empty, 0 when
// both are), summed in a fixed order: per-thread strided sums, the wavefront's sums, then the
// wavefronts' sums in order.  red holds 3 values per wavefront; every thread calls it (two barriers).
template <typename T>
__device__ __forceinline__ void score_origin(const T* __restrict__ A, int N, const T* __restrict__ Bp, int M,
                                             T* red, T c[3]) {
  const T* src = N > 0 ? A : Bp;
  const int n = N > 0 ? N : M;
  T sx = 0, sy = 0, sz = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sx += src[4 * (long)i];
    sy += src[4 * (long)i + 1];
    sz += src[4 * (long)i + 2];
  }
  sx = group_sum(sx, 64);
  sy = group_sum(sy, 64);
  sz = group_sum(sz, 64);
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[3 * wave] = sx;
    red[3 * wave + 1] = sy;
    red[3 * wave + 2] = sz;
  }
  __syncthreads();
  T t[3] = {0, 0, 0};
  for (int w = 0; w < nw; ++w)
    for (int k = 0; k < 3; ++k) t[k] += red[3 * w + k];
  __syncthreads();   // the caller reuses red
  for (int k = 0; k < 3; ++k) c[k] = n > 0 ? t[k] / (T)n : (T)0;
}

// Scale factors and origin of a pair's frame.
template <typename T, bool ESP>
__device__ __forceinline__ void score_frame(const T* __restrict__ A, int N, const T* __restrict__ Bp, int M,
                                            T alpha, T lam, T* red, T c[3], T& s, T& sw) {
  const T base = score_base<T>();
  s = sqrt(alpha / 2 * base);
  sw = ESP ? sqrt(base / lam) : (T)0;
  score_origin<T>(A, N, Bp, M, red, c);
}

// Scaled 4-D point i of a cloud.
template <typename T, bool ESP>
__device__ __forceinline__ typename Vec4<T>::type score_point(const T* __restrict__ src, int i, const T c[3],
                                                              T s, T sw) {
  const T* p = src + 4 * (long)i;
  return {(p[0] - c[0]) * s, (p[1] - c[1]) * s, (p[2] - c[2]) * s, ESP ? p[3] * sw : (T)0};
}

// W = -|p|^2 of a scaled point (the expanded form's per-point term).
__device__ __forceinline__ float neg_sq(const float4& p) {
  return -fmaf(p.x, p.x, fmaf(p.y, p.y, fmaf(p.z, p.z, p.w * p.w)));
}

// The expanded form's terms W_p, W_q and 2 p.q cancel to -|p - q|^2; their rounding grows with
// |p|^2.  A pair takes it only while every scaled point has |p|^2 <= kScoreHMax (about 30 A from
// the ref centroid at alpha 0.81); other pairs take the direct form.  Measured on 1024 surface +
// ESP pairs of 300 points (|p|^2 up to 227) against the FP64 kernels: S within 1.6e-7, dS/dfit
// within 4e-6 of its largest component, the same as the direct form's.
constexpr float kScoreHMax = 512.f;

// Whether a pair's clouds leave the expanded form's range (FP32; every thread calls it, two
// barriers; the same answer staged and streamed, from the same points).
template <typename T, bool ESP>
__device__ __forceinline__ bool score_far(const T* __restrict__ A, int N, const T* __restrict__ Bp, int M,
                                          const T c[3], T s, T sw, T* red) {
  if constexpr (!score_expanded<T>()) {
    return true;
  } else {
    float h = 0.f;
    for (int i = threadIdx.x; i < N + M; i += blockDim.x)
      h = fmaxf(h, -neg_sq(score_point<T, ESP>(i < N ? A : Bp, i < N ? i : i - N, c, s, sw)));
    for (int m = 32; m >= 1; m >>= 1) h = fmaxf(h, __shfl_xor(h, m, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = h;
    __syncthreads();
    float hm = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) hm = fmaxf(hm, red[w]);
    __syncthreads();   // the caller reuses red
    return !(hm <= kScoreHMax);   // fmaxf skips NaN points (their sums are NaN in either form)
  }
}

// Stages the pair's clouds (ref then fit) in the frame of score_frame, with W (FP32) after them,
// in one pass over HBM: the raw points go to LDS while the origin's sums run (the same threads
// summing the same points in the same order as score_origin), then they are moved into the frame
// in place, with score_far's test on the way (a maximum: the same in any order).
template <typename T, bool ESP>
__device__ __forceinline__ void stage_pair_clouds(const T* __restrict__ A, int N, const T* __restrict__ Bp, int M,
                                                  T alpha, T lam, T* red, typename Vec4<T>::type* la,
                                                  typename Vec4<T>::type* lb, T* wa, T* wb, T& s, bool& far) {
  using V = typename Vec4<T>::type;
  const int n0 = N > 0 ? N : M;   // the cloud whose centroid is the origin comes first
  T sx = 0, sy = 0, sz = 0;
  for (int i = threadIdx.x; i < N + M; i += blockDim.x) {
    const bool is_ref = i < N;
    const T* q = (is_ref ? A : Bp) + 4 * (long)(is_ref ? i : i - N);
    const V raw = {q[0], q[1], q[2], q[3]};
    (is_ref ? la[i] : lb[i - N]) = raw;
    if (i < n0) {
      sx += raw.x;
      sy += raw.y;
      sz += raw.z;
    }
  }
  sx = group_sum(sx, 64);
  sy = group_sum(sy, 64);
  sz = group_sum(sz, 64);
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[3 * wave] = sx;
    red[3 * wave + 1] = sy;
    red[3 * wave + 2] = sz;
  }
  __syncthreads();   // also: the raw points are in LDS
  T t[3] = {0, 0, 0};
  for (int w = 0; w < nw; ++w)
    for (int k = 0; k < 3; ++k) t[k] += red[3 * w + k];
  T c[3];
  for (int k = 0; k < 3; ++k) c[k] = n0 > 0 ? t[k] / (T)n0 : (T)0;
  const T base = score_base<T>();
  s = sqrt(alpha / 2 * base);
  const T sw = ESP ? sqrt(base / lam) : (T)0;
  float h = 0.f;
  for (int i = threadIdx.x; i < N + M; i += blockDim.x) {
    const bool is_ref = i < N;
    V& slot = is_ref ? la[i] : lb[i - N];
    const V raw = slot;
    const V p = {(raw.x - c[0]) * s, (raw.y - c[1]) * s, (raw.z - c[2]) * s, ESP ? raw.w * sw : (T)0};
    slot = p;
    if constexpr (score_expanded<T>()) {
      const float w = neg_sq(p);
      (is_ref ? wa[i] : wb[i - N]) = w;
      h = fmaxf(h, -w);
    }
  }
  if constexpr (score_expanded<T>()) {
    for (int m = 32; m >= 1; m >>= 1) h = fmaxf(h, __shfl_xor(h, m, 64));
    __syncthreads();   // every thread has read the origin's sums from red
    if ((threadIdx.x & 63) == 0) red[wave] = h;
    __syncthreads();   // also: the staged points are in LDS
    float hm = 0.f;
    for (int w = 0; w < nw; ++w) hm = fmaxf(hm, red[w]);
    far = !(hm <= kScoreHMax);
    __syncthreads();   // the caller reuses red
  } else {
    far = true;
    __syncthreads();
  }
}

// base^(-|q - p|^2) of one (column, row) point pair, direct form.
template <typename T, bool ESP>
__device__ __forceinline__ T pair_exp(const typename Vec4<T>::type& q, const typename Vec4<T>::type& p) {
  const T dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
  T r2 = dx * dx + dy * dy + dz * dz;
  if (ESP) { const T dw = q.w - p.w; r2 += dw * dw; }
  return neg_exp<T>(r2);
}

// -|q - p|^2 in the expanded form: W_q + W_p + 2 p.q, with p2 = 2 p.
template <bool ESP>
__device__ __forceinline__ float expo(const float4& q, float wq, const float4& p2, float wp) {
  float g = wq + wp;
  if (ESP) g = fmaf(q.w, p2.w, g);
  g = fmaf(q.z, p2.z, g);
  g = fmaf(q.y, p2.y, g);
  return fmaf(q.x, p2.x, g);
}

// Sum over the columns [0, nc) of base^(-|q_c - p|^2); every lane reads the same column (LDS
// broadcast).  Four running sums take the columns c % 4 (four loads in flight per step), added
// in a fixed order at the end.  Direct form (FP64).
template <typename T, bool ESP>
__device__ __forceinline__ T tile_sum(const typename Vec4<T>::type* __restrict__ cols, int nc,
                                      typename Vec4<T>::type p) {
  using V = typename Vec4<T>::type;
  T e[4] = {0, 0, 0, 0};
  int c = 0;
  for (; c + 3 < nc; c += 4) {
    V q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = cols[c + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) e[k] += pair_exp<T, ESP>(q[k], p);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (c + k < nc) e[k] += pair_exp<T, ESP>(cols[c + k], p);
  return (e[0] + e[1]) + (e[2] + e[3]);
}

// tile_sum in the expanded form (FP32): the same columns in the same running sums.
template <bool ESP>
__device__ __forceinline__ float tile_sum_x(const float4* __restrict__ cols, const float* __restrict__ cw, int nc,
                                            const float4& p2, float wp) {
  float e[4] = {0.f, 0.f, 0.f, 0.f};
  int c = 0;
  for (; c + 3 < nc; c += 4) {
    float4 q[4];
    float w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      q[k] = cols[c + k];
      w[k] = cw[c + k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) e[k] += fast_exp2(expo<ESP>(q[k], w[k], p2, wp));
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (c + k < nc) e[k] += fast_exp2(expo<ESP>(cols[c + k], cw[c + k], p2, wp));
  return (e[0] + e[1]) + (e[2] + e[3]);
}

// One 64-column tile of a row against the tile's columns, either form.
template <typename T, bool ESP>
__device__ __forceinline__ T row_tile(const typename Vec4<T>::type* __restrict__ cols, const T* __restrict__ cw,
                                      int nc, const typename Vec4<T>::type& p, T wp, bool far) {
  if constexpr (score_expanded<T>()) {
    if (!far) {
      const float4 p2 = {2.f * p.x, 2.f * p.y, 2.f * p.z, 2.f * p.w};
      return tile_sum_x<ESP>(cols, cw, nc, p2, wp);
    }
  }
  return tile_sum<T, ESP>(cols, nc, p);
}

// Forward: VAB, VAA and VBB as 64 x 64 tiles (lane = row, columns broadcast from LDS), VAA and VBB
// over the upper block triangle only (off-diagonal tiles count twice).  The tiles are dealt to the
// wavefronts round-robin, so a pair's whole product is balanced over its workgroup.
template <typename T, bool ESP>
__global__ __launch_bounds__(kScoreBlock) void k_score_fwd(const T* __restrict__ ref, const T* __restrict__ fit,
                                                          const int32_t* __restrict__ rr,
                                                          const int32_t* __restrict__ fr, T alpha, T lam,
                                                          T* S, T* ov) {
  using V = typename Vec4<T>::type;
  extern __shared__ double4 score_smem[];
  __shared__ T red[3 * (kScoreBlock / 64)];
  const int pair = blockIdx.x;
  const int N = rr[2 * pair + 1], M = fr[2 * pair + 1];
  V* la = reinterpret_cast<V*>(score_smem);
  V* lb = la + N;
  T* wa = reinterpret_cast<T*>(lb + M);   // FP32 only
  T* wb = wa + N;
  T s;
  bool far;
  stage_pair_clouds<T, ESP>(ref + 4 * (long)rr[2 * pair], N, fit + 4 * (long)fr[2 * pair], M, alpha, lam, red,
                            la, lb, wa, wb, s, far);
  constexpr int nw = kScoreBlock / 64;   // wavefronts (the launch's block is kScoreBlock)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  T acc[3] = {0, 0, 0};   // VAB, VAA, VBB sums
  int u = 0;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const V* rows = t == 1 ? la : lb;
    const V* cols = t == 2 ? lb : la;
    const T* rw = t == 1 ? wa : wb;
    const T* cw = t == 2 ? wb : wa;
    const int nr = t == 1 ? N : M, nc = t == 2 ? M : N;
    const int nbr = (nr + 63) / 64, nbc = (nc + 63) / 64;
    for (int i = 0; i < nbr; ++i) {
      const int r = i * 64 + lane, rc = r < nr ? r : nr - 1;
      const V p = rows[rc];
      const T wp = score_expanded<T>() ? rw[rc] : (T)0;
      // the row block's tiles j0 .. nbc - 1 are units u .. u + nbc - j0 - 1; this wavefront's
      // are those with unit % nw == wave
      const int j0 = t == 0 ? 0 : i;
      for (int j = j0 + ((wave - u) & (nw - 1)); j < nbc; j += nw) {
        const T e = row_tile<T, ESP>(cols + 64 * j, cw + 64 * j, min(64, nc - 64 * j), p, wp, far);
        acc[t] += r < nr ? (t != 0 && j > i ? 2 * e : e) : (T)0;
      }
      u += nbc - j0;
    }
  }
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const T v = group_sum(acc[t], 64);
    if (lane == 0) red[3 * wave + t] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    T v[3] = {0, 0, 0};
    for (int w = 0; w < nw; ++w)
      for (int t = 0; t < 3; ++t) v[t] += red[3 * w + t];
    const T K = (T)pow(M_PI / (2.0 * (double)alpha), 1.5);
    const T vab = v[0] * K, vaa = v[1] * K, vbb = v[2] * K;
    S[pair] = vab / (vaa + vbb - vab);
    ov[4 * pair] = vab; ov[4 * pair + 1] = vaa; ov[4 * pair + 2] = vbb; ov[4 * pair + 3] = 0;
  }
}

// Sum over the columns first, first + kScoreSlices, ... < n of base^(-|q - p|^2) (q - p) into two
// running sums, a over the slice's even-numbered columns and b over its odd-numbered ones (two
// loads in flight); the caller combines them after its last call (slice_total).  A block of 64
// columns holds an even number of each slice's columns, so calls over consecutive 64-column
// blocks add the same terms to the same sums in the same order as one call over the whole range.
// Direct form (FP64): a[0..2] are the force sums.
template <typename T, bool ESP>
__device__ __forceinline__ void slice_force(const typename Vec4<T>::type* __restrict__ cols, int n, int first,
                                            typename Vec4<T>::type p, T a[4], T b[4]) {
  using V = typename Vec4<T>::type;
  int c = first;
  for (; c + kScoreSlices < n; c += 2 * kScoreSlices) {
    const V q0 = cols[c], q1 = cols[c + kScoreSlices];
    const T dx0 = q0.x - p.x, dy0 = q0.y - p.y, dz0 = q0.z - p.z;
    const T dx1 = q1.x - p.x, dy1 = q1.y - p.y, dz1 = q1.z - p.z;
    T r0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0, r1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
    if (ESP) {
      const T dw0 = q0.w - p.w, dw1 = q1.w - p.w;
      r0 += dw0 * dw0;
      r1 += dw1 * dw1;
    }
    const T e0 = neg_exp<T>(r0), e1 = neg_exp<T>(r1);
    a[0] += e0 * dx0; a[1] += e0 * dy0; a[2] += e0 * dz0;
    b[0] += e1 * dx1; b[1] += e1 * dy1; b[2] += e1 * dz1;
  }
  if (c < n) {
    const V q = cols[c];
    const T dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
    T r2 = dx * dx + dy * dy + dz * dz;
    if (ESP) { const T dw = q.w - p.w; r2 += dw * dw; }
    const T e = neg_exp<T>(r2);
    a[0] += e * dx; a[1] += e * dy; a[2] += e * dz;
  }
}

// slice_force in the expanded form (FP32): a, b = (sum E, sum E q), the force sum E (q - p) formed
// once at the end (slice_total).
template <bool ESP>
__device__ __forceinline__ void slice_force_x(const float4* __restrict__ cols, const float* __restrict__ cw, int n,
                                              int first, const float4& p2, float wp, float a[4], float b[4]) {
  int c = first;
  // four columns per step (a, b, a, b): the same order per running sum as two per step
  for (; c + 3 * kScoreSlices < n; c += 4 * kScoreSlices) {
    float4 q[4];
    float w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      q[k] = cols[c + k * kScoreSlices];
      w[k] = cw[c + k * kScoreSlices];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float e = fast_exp2(expo<ESP>(q[k], w[k], p2, wp));
      float* r = (k & 1) ? b : a;
      r[0] += e; r[1] = fmaf(e, q[k].x, r[1]); r[2] = fmaf(e, q[k].y, r[2]); r[3] = fmaf(e, q[k].z, r[3]);
    }
  }
  for (; c + kScoreSlices < n; c += 2 * kScoreSlices) {
    const float4 q0 = cols[c], q1 = cols[c + kScoreSlices];
    const float w0 = cw[c], w1 = cw[c + kScoreSlices];
    const float e0 = fast_exp2(expo<ESP>(q0, w0, p2, wp)), e1 = fast_exp2(expo<ESP>(q1, w1, p2, wp));
    a[0] += e0; a[1] = fmaf(e0, q0.x, a[1]); a[2] = fmaf(e0, q0.y, a[2]); a[3] = fmaf(e0, q0.z, a[3]);
    b[0] += e1; b[1] = fmaf(e1, q1.x, b[1]); b[2] = fmaf(e1, q1.y, b[2]); b[3] = fmaf(e1, q1.z, b[3]);
  }
  if (c < n) {
    const float4 q = cols[c];
    const float e = fast_exp2(expo<ESP>(q, cw[c], p2, wp));
    a[0] += e; a[1] = fmaf(e, q.x, a[1]); a[2] = fmaf(e, q.y, a[2]); a[3] = fmaf(e, q.z, a[3]);
  }
}
static_assert(64 % (2 * kScoreSlices) == 0, "a 64-column block must hold an even count of each slice's columns");

// One call of either form over columns [0, n) of a block.
template <typename T, bool ESP>
__device__ __forceinline__ void row_slice(const typename Vec4<T>::type* __restrict__ cols, const T* __restrict__ cw,
                                          int n, int first, const typename Vec4<T>::type& p, T wp, bool far,
                                          T a[4], T b[4]) {
  if constexpr (score_expanded<T>()) {
    if (!far) {
      const float4 p2 = {2.f * p.x, 2.f * p.y, 2.f * p.z, 2.f * p.w};
      slice_force_x<ESP>(cols, cw, n, first, p2, wp, a, b);
      return;
    }
  }
  slice_force<T, ESP>(cols, n, first, p, a, b);
}

// The force sum E (q - p) of a row from its two running sums (expanded form: (E, E q) sums).
template <typename T>
__device__ __forceinline__ void slice_total(const T a[4], const T b[4], const typename Vec4<T>::type& p, bool far,
                                            T f[3]) {
  if (score_expanded<T>() && !far) {
    const T e = a[0] + b[0];
    f[0] = (a[1] + b[1]) - p.x * e;
    f[1] = (a[2] + b[2]) - p.y * e;
    f[2] = (a[3] + b[3]) - p.z * e;
  } else {
    for (int k = 0; k < 3; ++k) f[k] = a[k] + b[k];
  }
}

// Backward.  With U = VAA + VBB and D = U - VAB, for a fit point x_b
//   dS/dx_b = K alpha / D^2 (U sum_a w E (a - x_b) - 2 VAB sum_b' w E (x_b' - x_b))
// (the second sum is dVBB/dx_b / 2; VBB is invariant only under rigid motions), and the same
// with the clouds' roles swapped for a ref point.  A wavefront unit is kScoreRows rows; the
// kScoreSlices lanes of a row split its columns and are combined with a fixed shuffle order, so
// the result is deterministic.
template <typename T, bool ESP>
__global__ __launch_bounds__(kScoreBlock) void k_score_bwd(const T* __restrict__ ref, const T* __restrict__ fit,
                                                          const int32_t* __restrict__ rr,
                                                          const int32_t* __restrict__ fr, T alpha, T lam,
                                                          const T* __restrict__ ov, const T* __restrict__ gS,
                                                          T* d_fit, T* d_ref) {
  using V = typename Vec4<T>::type;
  extern __shared__ double4 score_smem[];
  __shared__ T red[3 * (kScoreBlock / 64)];
  const int pair = blockIdx.x;
  const int N = rr[2 * pair + 1], M = fr[2 * pair + 1];
  V* la = reinterpret_cast<V*>(score_smem);
  V* lb = la + N;
  T* wa = reinterpret_cast<T*>(lb + M);   // FP32 only
  T* wb = wa + N;
  T s;
  bool far;
  stage_pair_clouds<T, ESP>(ref + 4 * (long)rr[2 * pair], N, fit + 4 * (long)fr[2 * pair], M, alpha, lam, red,
                            la, lb, wa, wb, s, far);
  const T K = (T)pow(M_PI / (2.0 * (double)alpha), 1.5);
  const T vab = ov[4 * pair], U = ov[4 * pair + 1] + ov[4 * pair + 2], D = U - vab;
  const T g = gS[pair] * K * alpha / (D * D) / s;   // the force sums are in scaled coordinates
  constexpr int nw = kScoreBlock / 64;   // wavefronts (the launch's block is kScoreBlock)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sl = lane / kScoreRows, rl = lane % kScoreRows;
  const int nbf = (M + kScoreRows - 1) / kScoreRows, nba = d_ref ? (N + kScoreRows - 1) / kScoreRows : 0;
  for (int u = wave; u < nbf + nba; u += nw) {
    const bool fit_row = u < nbf;
    const V* own = fit_row ? lb : la;
    const V* oth = fit_row ? la : lb;
    const T* ownw = fit_row ? wb : wa;
    const T* othw = fit_row ? wa : wb;
    const int n_own = fit_row ? M : N, n_oth = fit_row ? N : M;
    const int row = (fit_row ? u : u - nbf) * kScoreRows + rl, rc = row < n_own ? row : n_own - 1;
    const V p = own[rc];
    const T wp = score_expanded<T>() ? ownw[rc] : (T)0;
    T oa[4] = {0, 0, 0, 0}, ob[4] = {0, 0, 0, 0}, sa[4] = {0, 0, 0, 0}, sb[4] = {0, 0, 0, 0};
    row_slice<T, ESP>(oth, othw, n_oth, sl, p, wp, far, oa, ob);
    row_slice<T, ESP>(own, ownw, n_own, sl, p, wp, far, sa, sb);
    T fo[3], fs[3];
    slice_total<T>(oa, ob, p, far, fo);
    slice_total<T>(sa, sb, p, far, fs);
    T d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T v = U * fo[k] - 2 * vab * fs[k];
      v += __shfl_xor(v, kScoreRows);
      v += __shfl_xor(v, 2 * kScoreRows);
      d[k] = g * v;
    }
    if (sl == 0 && row < n_own) {
      T* o = fit_row ? d_fit + 4 * ((long)fr[2 * pair] + row) : d_ref + 4 * ((long)rr[2 * pair] + row);
      o[0] = d[0]; o[1] = d[1]; o[2] = d[2]; o[3] = 0;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Clouds beyond LDS: the same forward and backward with each wavefront staging the 64 points of
// the column block it works on (one per lane) into a buffer of its own, and the row points read
// from HBM; the frame, the column partition per row (blocks of 64; slice c % 4 in the backward)
// and the per-point values are the staged kernels', so the sums are taken in the same order.
// ------------------------------------------------------------------------------------------------
// Points [c0, c0 + n) of a cloud into the calling wavefront's buffers (lane k writes point c0 + k
// and, FP32, its W).  The wavefront-scope fences order the stores after the previous block's loads
// and before this block's (a wavefront's LDS operations complete in issue order).
template <typename T, bool ESP>
__device__ __forceinline__ void stage_block(const T* __restrict__ src, int c0, int n, const T c[3], T s, T sw,
                                            typename Vec4<T>::type* wbuf, T* wbw) {
  const int lane = threadIdx.x & 63;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < n) {
    const typename Vec4<T>::type p = score_point<T, ESP>(src, c0 + lane, c, s, sw);
    wbuf[lane] = p;
    if constexpr (score_expanded<T>()) wbw[lane] = neg_sq(p);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// W of a row point read from HBM (FP32; 0 for the direct form).
template <typename T>
__device__ __forceinline__ T row_w(const typename Vec4<T>::type& p) {
  if constexpr (score_expanded<T>()) return neg_sq(p);
  else return (T)0;
}

template <typename T, bool ESP>
__global__ __launch_bounds__(kScoreBlock) void k_score_fwd_stream(const T* __restrict__ ref, const T* __restrict__ fit,
                                                                 const int32_t* __restrict__ rr,
                                                                 const int32_t* __restrict__ fr, T alpha, T lam,
                                                                 T* S, T* ov) {
  using V = typename Vec4<T>::type;
  __shared__ V wbufs[kScoreBlock];
  __shared__ T wbws[kScoreBlock];
  __shared__ T red[3 * (kScoreBlock / 64)];
  const int pair = blockIdx.x;
  const int N = rr[2 * pair + 1], M = fr[2 * pair + 1];
  const T* A = ref + 4 * (long)rr[2 * pair];
  const T* Bp = fit + 4 * (long)fr[2 * pair];
  T c[3], s, sw;
  score_frame<T, ESP>(A, N, Bp, M, alpha, lam, red, c, s, sw);
  const bool far = score_far<T, ESP>(A, N, Bp, M, c, s, sw, red);
  constexpr int nw = kScoreBlock / 64;   // wavefronts (the launch's block is kScoreBlock)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  V* wbuf = wbufs + 64 * wave;
  T* wbw = wbws + 64 * wave;
  T acc[3] = {0, 0, 0};
  int u = 0;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const T* rows = t == 1 ? A : Bp;
    const T* cols = t == 2 ? Bp : A;
    const int nr = t == 1 ? N : M, nc = t == 2 ? M : N;
    const int nbr = (nr + 63) / 64, nbc = (nc + 63) / 64;
    for (int i = 0; i < nbr; ++i) {
      const int r = i * 64 + lane;
      const V p = score_point<T, ESP>(rows, r < nr ? r : nr - 1, c, s, sw);
      const T wp = row_w<T>(p);
      const int j0 = t == 0 ? 0 : i;
      for (int j = j0 + ((wave - u) & (nw - 1)); j < nbc; j += nw) {
        const int n = min(64, nc - 64 * j);
        stage_block<T, ESP>(cols, 64 * j, n, c, s, sw, wbuf, wbw);
        const T e = row_tile<T, ESP>(wbuf, wbw, n, p, wp, far);
        acc[t] += r < nr ? (t != 0 && j > i ? 2 * e : e) : (T)0;
      }
      u += nbc - j0;
    }
  }
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const T v = group_sum(acc[t], 64);
    if (lane == 0) red[3 * wave + t] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    T v[3] = {0, 0, 0};
    for (int w = 0; w < nw; ++w)
      for (int t = 0; t < 3; ++t) v[t] += red[3 * w + t];
    const T K = (T)pow(M_PI / (2.0 * (double)alpha), 1.5);
    const T vab = v[0] * K, vaa = v[1] * K, vbb = v[2] * K;
    S[pair] = vab / (vaa + vbb - vab);
    ov[4 * pair] = vab; ov[4 * pair + 1] = vaa; ov[4 * pair + 2] = vbb; ov[4 * pair + 3] = 0;
  }
}

template <typename T, bool ESP>
__global__ __launch_bounds__(kScoreBlock) void k_score_bwd_stream(const T* __restrict__ ref, const T* __restrict__ fit,
                                                                 const int32_t* __restrict__ rr,
                                                                 const int32_t* __restrict__ fr, T alpha, T lam,
                                                                 const T* __restrict__ ov, const T* __restrict__ gS,
                                                                 T* d_fit, T* d_ref) {
  using V = typename Vec4<T>::type;
  __shared__ V wbufs[kScoreBlock];
  __shared__ T wbws[kScoreBlock];
  __shared__ T red[3 * (kScoreBlock / 64)];
  const int pair = blockIdx.x;
  const int N = rr[2 * pair + 1], M = fr[2 * pair + 1];
  const T* A = ref + 4 * (long)rr[2 * pair];
  const T* Bp = fit + 4 * (long)fr[2 * pair];
  T c[3], s, sw;
  score_frame<T, ESP>(A, N, Bp, M, alpha, lam, red, c, s, sw);
  const bool far = score_far<T, ESP>(A, N, Bp, M, c, s, sw, red);
  const T K = (T)pow(M_PI / (2.0 * (double)alpha), 1.5);
  const T vab = ov[4 * pair], U = ov[4 * pair + 1] + ov[4 * pair + 2], D = U - vab;
  const T g = gS[pair] * K * alpha / (D * D) / s;
  constexpr int nw = kScoreBlock / 64;   // wavefronts (the launch's block is kScoreBlock)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  V* wbuf = wbufs + 64 * wave;
  T* wbw = wbws + 64 * wave;
  const int sl = lane / kScoreRows, rl = lane % kScoreRows;
  const int nbf = (M + kScoreRows - 1) / kScoreRows, nba = d_ref ? (N + kScoreRows - 1) / kScoreRows : 0;
  for (int u = wave; u < nbf + nba; u += nw) {
    const bool fit_row = u < nbf;
    const T* own = fit_row ? Bp : A;
    const T* oth = fit_row ? A : Bp;
    const int n_own = fit_row ? M : N, n_oth = fit_row ? N : M;
    const int row = (fit_row ? u : u - nbf) * kScoreRows + rl;
    const V p = score_point<T, ESP>(own, row < n_own ? row : n_own - 1, c, s, sw);
    const T wp = row_w<T>(p);
    T oa[4] = {0, 0, 0, 0}, ob[4] = {0, 0, 0, 0}, sa[4] = {0, 0, 0, 0}, sb[4] = {0, 0, 0, 0};
    for (int c0 = 0; c0 < n_oth; c0 += 64) {
      const int n = min(64, n_oth - c0);
      stage_block<T, ESP>(oth, c0, n, c, s, sw, wbuf, wbw);
      row_slice<T, ESP>(wbuf, wbw, n, sl, p, wp, far, oa, ob);
    }
    for (int c0 = 0; c0 < n_own; c0 += 64) {
      const int n = min(64, n_own - c0);
      stage_block<T, ESP>(own, c0, n, c, s, sw, wbuf, wbw);
      row_slice<T, ESP>(wbuf, wbw, n, sl, p, wp, far, sa, sb);
    }
    T fo[3], fs[3];
    slice_total<T>(oa, ob, p, far, fo);
    slice_total<T>(sa, sb, p, far, fs);
    T d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T v = U * fo[k] - 2 * vab * fs[k];
      v += __shfl_xor(v, kScoreRows);
      v += __shfl_xor(v, 2 * kScoreRows);
      d[k] = g * v;
    }
    if (sl == 0 && row < n_own) {
      T* o = fit_row ? d_fit + 4 * ((long)fr[2 * pair] + row) : d_ref + 4 * ((long)rr[2 * pair] + row);
      o[0] = d[0]; o[1] = d[1]; o[2] = d[2]; o[3] = 0;
    }
  }
}

template <int MODE>
void launch_self(const FssPairBatch* pb, const KScore& ks, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_self_overlap<MODE>, dim3(pb->n_pairs, 2), dim3(kPoseBlock), 0, s, *pb, ks, out);
}

// Staged clouds whose fit rows (shape and pharmacophore) fit one wavefront run one wavefront per
// pose: a lane owns the same fit rows as in four wavefronts, whose other lanes would only add
// zeros, so the sums and their bits are the same and the pose needs no exchange across
// wavefronts.  (The staging centroid is summed independently of the workgroup size.)
template <int MODE, bool GRAD, int OUT>
void launch_pose(const FssPairBatch* pb, const KScore& ks, LdsPlan lp, int n, const int32_t* pp,
                 const float* q, const float* t, const double* so, float invB, float* V, float* gR,
                 float* gT, float* gQ, float* loss, float* grad, hipStream_t s) {
  const bool one_wave = MODE != FSS_MODE_ESP_COMBO && lp.tile_ref == 0 && pb->max_fit <= 64 &&
                        pb->max_fit_ph <= 64;
  hipLaunchKernelGGL((k_pose_grad<MODE, GRAD, OUT>), dim3(n), dim3(one_wave ? 64 : kPoseBlock),
                     lp.end * sizeof(float4), s, *pb, ks, lp, pp, q, t, so, invB, V, gR, gT, gQ, loss, grad);
}

// Workgroups of up to 512 threads get a 256-VGPR budget, larger ones 128.  Shape inner loop:
// variant 1 = direct form, 3 = split-bf16 MFMA (one wavefront per pose), otherwise the expanded
// form; the pharmacophore-only mode has no shape loop and always runs as variant 2.
int effective_variant(int mode, int variant) {
  if (mode == FSS_MODE_PHARM) return 2;
  if (mode == FSS_MODE_ESP_COMBO) return variant == 1 ? 1 : 2;
  if (variant == 4) return tileable(mode) ? 4 : 2;
  return variant == 1 || variant == 3 ? variant : 2;
}
template <int MODE, bool ES>
void launch_align(const FssPairBatch* pb, const KScore& ks, LdsPlan lp, const AlignK& ak, int grid,
                  int block, size_t lds, int variant, hipStream_t s) {
  const int v = effective_variant(MODE, variant);
#define FSS_K(T, V) hipLaunchKernelGGL((k_align<MODE, ES, T, V>), dim3(grid), dim3(block), lds, s, *pb, ks, lp, ak)
  if constexpr (tileable(MODE)) {
    if (v == 4) {   // streamed shape clouds
      if (block <= 512) FSS_K(512, 4);
      else FSS_K(1024, 4);
      return;
    }
  }
  if (ak.ppw > 1) {   // several pairs per workgroup: VALU inner loops, at most 512 lanes
    if (v == 1) hipLaunchKernelGGL((k_align<MODE, ES, 512, 1, true>), dim3(grid), dim3(block), lds, s, *pb, ks, lp, ak);
    else hipLaunchKernelGGL((k_align<MODE, ES, 512, 2, true>), dim3(grid), dim3(block), lds, s, *pb, ks, lp, ak);
  } else if (block <= 512) {
    if (v == 1) FSS_K(512, 1);
    else if (v == 3) FSS_K(512, 3);
    else FSS_K(512, 2);
  } else {
    if (v == 1) FSS_K(1024, 1);
    else if (v == 3) FSS_K(1024, 3);
    else FSS_K(1024, 2);
  }
#undef FSS_K
}

int score_check(const FssPairBatch* batch, const FssScoreParams* sp) {
  if (!sp) return set_error(FSS_E_INVALID, "score params are NULL");
  if (sp->mode != FSS_MODE_SHAPE && sp->mode != FSS_MODE_SHAPE_ESP)
    return set_error(FSS_E_INVALID, "score fwd/bwd supports the shape and ESP modes only");
  if (!(sp->alpha > 0.0)) return set_error(FSS_E_INVALID, "alpha must be > 0");
  if (sp->mode == FSS_MODE_SHAPE_ESP && !(sp->lam > 0.0)) return set_error(FSS_E_INVALID, "lam must be > 0");
  if (!batch || !batch->ref_rng || !batch->fit_rng) return set_error(FSS_E_INVALID, "NULL ranges");
  return FSS_OK;
}

// Both clouds of a pair are staged in LDS (4 values per point, and W in FP32) when they fit beside
// the static reduction scratch (3 sums per wavefront); lds = 0 selects the streamed kernels (a
// 64-point block per wavefront).
int score_lds(const FssPairBatch* batch, size_t elem, size_t& lds) {
  if (batch->max_ref < 0 || batch->max_fit < 0) return set_error(FSS_E_INVALID, "negative max_ref / max_fit");
  // FP32: the float4 point and its W; FP64: the double4 point
  lds = ((size_t)batch->max_ref + (size_t)batch->max_fit) * (elem == 4 ? 5 : 4) * elem;
  if (lds + 3 * (kScoreBlock / 64) * elem > max_lds_bytes()) lds = 0;
  return FSS_OK;
}

template <typename T>
int score_fwd_t(const FssPairBatch* batch, const FssScoreParams* sp, const T* ref, const T* fit,
                       T* S, T* ov, hipStream_t s) {
  int rc = score_check(batch, sp);
  if (rc) return rc;
  if (!ref || !fit || !S || !ov) return set_error(FSS_E_INVALID, "NULL score arrays");
  if (batch->n_pairs == 0) return FSS_OK;
  const T lam = sp->mode == FSS_MODE_SHAPE_ESP ? (T)sp->lam : (T)1;
  size_t lds = 0;
  if ((rc = score_lds(batch, sizeof(T), lds))) return rc;
  const bool esp = sp->mode == FSS_MODE_SHAPE_ESP;
  if (!lds && esp)
    hipLaunchKernelGGL((k_score_fwd_stream<T, true>), dim3(batch->n_pairs), dim3(kScoreBlock), 0, s, ref, fit, batch->ref_rng, batch->fit_rng, (T)sp->alpha, lam, S, ov);
  else if (!lds)
    hipLaunchKernelGGL((k_score_fwd_stream<T, false>), dim3(batch->n_pairs), dim3(kScoreBlock), 0, s, ref, fit, batch->ref_rng, batch->fit_rng, (T)sp->alpha, lam, S, ov);
  else if (esp)
    hipLaunchKernelGGL((k_score_fwd<T, true>), dim3(batch->n_pairs), dim3(kScoreBlock), lds, s, ref, fit, batch->ref_rng, batch->fit_rng, (T)sp->alpha, lam, S, ov);
  else
    hipLaunchKernelGGL((k_score_fwd<T, false>), dim3(batch->n_pairs), dim3(kScoreBlock), lds, s, ref, fit, batch->ref_rng, batch->fit_rng, (T)sp->alpha, lam, S, ov);
  return check_launch("fss_score_fwd");
}

template <typename T>
int score_bwd_t(const FssPairBatch* batch, const FssScoreParams* sp, const T* ref, const T* fit,
                       const T* ov, const T* gS, T* d_fit, T* d_ref, hipStream_t s) {
  int rc = score_check(batch, sp);
  if (rc) return rc;
  if (!ref || !fit || !ov || !gS || !d_fit) return set_error(FSS_E_INVALID, "NULL score arrays");
  if (batch->n_pairs == 0) return FSS_OK;
  const T lam = sp->mode == FSS_MODE_SHAPE_ESP ? (T)sp->lam : (T)1;
  size_t lds = 0;
  if ((rc = score_lds(batch, sizeof(T), lds))) return rc;
  const bool esp = sp->mode == FSS_MODE_SHAPE_ESP;
  if (!lds && esp)
    hipLaunchKernelGGL((k_score_bwd_stream<T, true>), dim3(batch->n_pairs), dim3(kScoreBlock), 0, s, ref, fit, batch->ref_rng, batch->fit_rng, (T)sp->alpha, lam, ov, gS, d_fit, d_ref);
  else if (!lds)
    hipLaunchKernelGGL((k_score_bwd_stream<T, false>), dim3(batch->n_pairs), dim3(kScoreBlock), 0, s, ref, fit, batch->ref_rng, batch->fit_rng, (T)sp->alpha, lam, ov, gS, d_fit, d_ref);
  else if (esp)
    hipLaunchKernelGGL((k_score_bwd<T, true>), dim3(batch->n_pairs), dim3(kScoreBlock), lds, s, ref, fit, batch->ref_rng, batch->fit_rng, (T)sp->alpha, lam, ov, gS, d_fit, d_ref);
  else
    hipLaunchKernelGGL((k_score_bwd<T, false>), dim3(batch->n_pairs), dim3(kScoreBlock), lds, s, ref, fit, batch->ref_rng, batch->fit_rng, (T)sp->alpha, lam, ov, gS, d_fit, d_ref);
  return check_launch("fss_score_bwd");
}

template <int MODE>
void launch_pose_score(const FssPairBatch* pb, const KScore& ks, int n, int B, const int32_t* pp,
                              const double* so, const float* q, const float* t, double* out, hipStream_t s,
                              const double* thr = nullptr) {
  hipLaunchKernelGGL(k_pose_score<MODE>, dim3((n + 3) / 4), dim3(256), 0, s, *pb, ks, n, B, pp, so, q, t, out, thr);
}

int pose_score_impl(const FssPairBatch* batch, const KScore& ks, int32_t n_poses,
                           const int32_t* pose_pair, int32_t B, const double* self_ov, const float* q,
                           const float* t, double* score, hipStream_t s) {
  switch (ks.mode) {
    case FSS_MODE_SHAPE: launch_pose_score<FSS_MODE_SHAPE>(batch, ks, n_poses, B, pose_pair, self_ov, q, t, score, s); break;
    case FSS_MODE_SHAPE_ESP: launch_pose_score<FSS_MODE_SHAPE_ESP>(batch, ks, n_poses, B, pose_pair, self_ov, q, t, score, s); break;
    case FSS_MODE_PHARM: launch_pose_score<FSS_MODE_PHARM>(batch, ks, n_poses, B, pose_pair, self_ov, q, t, score, s); break;
    case FSS_MODE_ESP_COMBO: launch_pose_score<FSS_MODE_ESP_COMBO>(batch, ks, n_poses, B, pose_pair, self_ov, q, t, score, s); break;
    default: launch_pose_score<FSS_MODE_SHAPE_PHARM>(batch, ks, n_poses, B, pose_pair, self_ov, q, t, score, s); break;
  }
  return check_launch("fss_pose_score");
}

// --- device seeds -------------------------------------------------------------------------------
// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (FP64): a -> diag, V columns.
__device__ void jacobi3(double a[3][3], double V[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) V[i][k] = i == k ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 24; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-30 * dia || off == 0.0) break;
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      if (a[p][q] == 0.0) continue;
      const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
      const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; ++k) {
        const double akp = a[k][p], akq = a[k][q];
        a[k][p] = c * akp - s * akq; a[k][q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; ++k) {
        const double apk = a[p][k], aqk = a[q][k];
        a[p][k] = c * apk - s * aqk; a[q][k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; ++k) {
        const double vkp = V[k][p], vkq = V[k][q];
        V[k][p] = c * vkp - s * vkq; V[k][q] = s * vkp + c * vkq;
      }
    }
  }
}

// Principal axes (rows, largest inertia eigenvalue first) from the centred second moments C:
// inertia = trace(C) I - C (pca_np.py:12-33).
__device__ void principal_axes(const double C[6], double ax[3][3]) {
  const double tr = C[0] + C[3] + C[5];
  double a[3][3] = {{tr - C[0], -C[1], -C[2]}, {-C[1], tr - C[3], -C[4]}, {-C[2], -C[4], tr - C[5]}};
  double V[3][3];
  jacobi3(a, V);
  double w[3] = {a[0][0], a[1][1], a[2][2]};
  int o[3] = {0, 1, 2};
  for (int i = 0; i < 2; ++i)
    for (int k = 0; k < 2 - i; ++k)
      if (w[o[k]] < w[o[k + 1]]) { const int tmp = o[k]; o[k] = o[k + 1]; o[k + 1] = tmp; }
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) ax[r][k] = V[k][o[r]];
}

// Quaternion rotating v1 onto v2 (axis-angle construction of pca_np.py:36-88).
__device__ void quat_between(const double v1[3], const double v2[3], double q[4]) {
  const double n1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  const double n2 = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  double d = (v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]) / (n1 * n2);
  d = fmin(fmax(d, -1.0), 1.0);
  const double ang = acos(d);
  double ax[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
  double na = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
  if (na < 1e-12) {            // parallel: any axis; antiparallel: an axis perpendicular to v1
    const bool xmin = fabs(v1[0]) <= fabs(v1[1]) && fabs(v1[0]) <= fabs(v1[2]);
    const double e[3] = {xmin ? 1.0 : 0.0, xmin ? 0.0 : 1.0, 0.0};
    ax[0] = v1[1] * e[2] - v1[2] * e[1]; ax[1] = v1[2] * e[0] - v1[0] * e[2]; ax[2] = v1[0] * e[1] - v1[1] * e[0];
    na = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
  }
  const double sn = sin(0.5 * ang) / na;
  q[0] = cos(0.5 * ang); q[1] = ax[0] * sn; q[2] = ax[1] * sn; q[3] = ax[2] * sn;
}

__device__ void qmul(const double p[4], const double q[4], double o[4]) {
  o[0] = p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3];
  o[1] = p[1] * q[0] + p[0] * q[1] - p[3] * q[2] + p[2] * q[3];
  o[2] = p[2] * q[0] + p[3] * q[1] + p[0] * q[2] - p[1] * q[3];
  o[3] = p[3] * q[0] - p[2] * q[1] + p[1] * q[2] + p[0] * q[3];
}

// Restart seeds, one wave per pair.  flags bit 0: use the pharmacophore anchors instead of the
// shape clouds; bit 1: the reference's seed rule (alignment/_torch.py:253-326): B == 1 is the
// identity with zero translation, rows 1-4 are the principal-axis alignments (B >= 5).  Other
// rows take table[r]; every translation is com(ref) - R(q) com(fit).
__global__ __launch_bounds__(64) void k_seed_poses(FssPairBatch pb, int flags, int B,
                                                   const float* __restrict__ table, float* q, float* t) {
  const int pair = blockIdx.x, j = threadIdx.x;
  const bool use_pharm = flags & 1, ref_rule = flags & 2;
  const int* rr = use_pharm ? pb.ref_prng : pb.ref_rng;
  const int* fr = use_pharm ? pb.fit_prng : pb.fit_rng;
  const float4* rp = reinterpret_cast<const float4*>(use_pharm ? pb.ref_anc : pb.ref_pts) + rr[2 * pair];
  const float4* fp = reinterpret_cast<const float4*>(use_pharm ? pb.fit_anc : pb.fit_pts) + fr[2 * pair];
  const int N = rr[2 * pair + 1], M = fr[2 * pair + 1];
  // first and second moments of both clouds: c[0..2], s[0..5] ref; c[3..5], s[6..11] fit
  double m[18];
  for (int k = 0; k < 18; ++k) m[k] = 0.0;
  for (int i = j; i < N; i += 64) {
    const double x = rp[i].x, y = rp[i].y, z = rp[i].z;
    m[0] += x; m[1] += y; m[2] += z;
    m[6] += x * x; m[7] += x * y; m[8] += x * z; m[9] += y * y; m[10] += y * z; m[11] += z * z;
  }
  for (int i = j; i < M; i += 64) {
    const double x = fp[i].x, y = fp[i].y, z = fp[i].z;
    m[3] += x; m[4] += y; m[5] += z;
    m[12] += x * x; m[13] += x * y; m[14] += x * z; m[15] += y * y; m[16] += y * z; m[17] += z * z;
  }
  for (int k = 0; k < 18; ++k) m[k] = group_sum(m[k], 64);
  const double iN = N > 0 ? 1.0 / N : 0.0, iM = M > 0 ? 1.0 / M : 0.0;
  const double cr[3] = {m[0] * iN, m[1] * iN, m[2] * iN}, cf[3] = {m[3] * iM, m[4] * iM, m[5] * iM};
  double pca[4][4];
  const bool do_pca = ref_rule && B >= 5;
  if (do_pca) {
    const double Cr[6] = {m[6] * iN - cr[0] * cr[0], m[7] * iN - cr[0] * cr[1], m[8] * iN - cr[0] * cr[2],
                          m[9] * iN - cr[1] * cr[1], m[10] * iN - cr[1] * cr[2], m[11] * iN - cr[2] * cr[2]};
    const double Cf[6] = {m[12] * iM - cf[0] * cf[0], m[13] * iM - cf[0] * cf[1], m[14] * iM - cf[0] * cf[2],
                          m[15] * iM - cf[1] * cf[1], m[16] * iM - cf[1] * cf[2], m[17] * iM - cf[2] * cf[2]};
    double ar[3][3], af[3][3];
    principal_axes(Cr, ar);
    principal_axes(Cf, af);
    const double fl[4][2] = {{1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
    for (int k = 0; k < 4; ++k) {
      const double t0[3] = {fl[k][0] * ar[0][0], fl[k][0] * ar[0][1], fl[k][0] * ar[0][2]};
      const double t1[3] = {fl[k][1] * ar[1][0], fl[k][1] * ar[1][1], fl[k][1] * ar[1][2]};
      double q1[4], q2[4], R1[9], qh[4], qn;
      quat_between(af[0], t0, q1);
      const float q1f[4] = {(float)q1[0], (float)q1[1], (float)q1[2], (float)q1[3]};
      rotation_fp64(q1f, qh, qn, R1);
      const double g1[3] = {R1[0] * af[1][0] + R1[1] * af[1][1] + R1[2] * af[1][2],
                            R1[3] * af[1][0] + R1[4] * af[1][1] + R1[5] * af[1][2],
                            R1[6] * af[1][0] + R1[7] * af[1][1] + R1[8] * af[1][2]};
      quat_between(g1, t1, q2);
      qmul(q2, q1, pca[k]);
    }
  }
  for (int r = j; r < B; r += 64) {
    float qq[4] = {table[4 * r], table[4 * r + 1], table[4 * r + 2], table[4 * r + 3]};
    if (do_pca && r >= 1 && r <= 4)
      for (int k = 0; k < 4; ++k) qq[k] = (float)pca[r - 1][k];
    if (ref_rule && r == 0) { qq[0] = 1.f; qq[1] = qq[2] = qq[3] = 0.f; }
    double qh[4], qn, R[9];
    rotation_fp64(qq, qh, qn, R);
    const long pose = (long)pair * B + r;
    for (int k = 0; k < 4; ++k) q[4 * pose + k] = qq[k];
    const bool zero_t = ref_rule && B == 1;
    for (int i = 0; i < 3; ++i)
      t[3 * pose + i] = zero_t ? 0.f : (float)(cr[i] - (R[3 * i] * cf[0] + R[3 * i + 1] * cf[1] + R[3 * i + 2] * cf[2]));
  }
}

// Final scores and the best restart of every pair.  Every mode but ShaEP: FP32 pre-pass over
// every final pose, per-pair threshold (kept in best_score until k_align_finish overwrites it),
// FP64 for the restarts at or above it.  ShaEP scores every final pose in FP64.
template <int MODE>
void launch_prepass(const FssPairBatch* batch, const KScore& ks, int n, int B, const double* self_ov,
                    FssAlignOut* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pose_score_f32<MODE>, dim3((n + 3) / 4), dim3(256), 0, s, *batch, ks, n, B, self_ov,
                     out->q, out->t, out->pose_score);
}

int align_finish_impl(const FssPairBatch* batch, const KScore& ks, int B, const double* self_ov,
                      FssAlignOut* out, hipStream_t s) {
  const int n = batch->n_pairs * B;
  int rc;
  if (ks.mode != FSS_MODE_ESP_COMBO) {
    switch (ks.mode) {
      case FSS_MODE_SHAPE: launch_prepass<FSS_MODE_SHAPE>(batch, ks, n, B, self_ov, out, s); break;
      case FSS_MODE_SHAPE_ESP: launch_prepass<FSS_MODE_SHAPE_ESP>(batch, ks, n, B, self_ov, out, s); break;
      case FSS_MODE_PHARM: launch_prepass<FSS_MODE_PHARM>(batch, ks, n, B, self_ov, out, s); break;
      default: launch_prepass<FSS_MODE_SHAPE_PHARM>(batch, ks, n, B, self_ov, out, s); break;
    }
    hipLaunchKernelGGL(k_rescore_threshold, dim3((batch->n_pairs + 255) / 256), dim3(256), 0, s, batch->n_pairs, B,
                       out->pose_score, out->best_score);
    if ((rc = check_launch("fss_align (re-scoring pre-pass)"))) return rc;
    switch (ks.mode) {
      case FSS_MODE_SHAPE:
        launch_pose_score<FSS_MODE_SHAPE>(batch, ks, n, B, nullptr, self_ov, out->q, out->t, out->pose_score, s,
                                          out->best_score);
        break;
      case FSS_MODE_SHAPE_ESP:
        launch_pose_score<FSS_MODE_SHAPE_ESP>(batch, ks, n, B, nullptr, self_ov, out->q, out->t, out->pose_score, s,
                                              out->best_score);
        break;
      case FSS_MODE_PHARM:
        launch_pose_score<FSS_MODE_PHARM>(batch, ks, n, B, nullptr, self_ov, out->q, out->t, out->pose_score, s,
                                          out->best_score);
        break;
      default:
        launch_pose_score<FSS_MODE_SHAPE_PHARM>(batch, ks, n, B, nullptr, self_ov, out->q, out->t, out->pose_score, s,
                                                out->best_score);
        break;
    }
    rc = check_launch("fss_align (FP64 re-scoring)");
  } else {
    rc = pose_score_impl(batch, ks, n, nullptr, B, self_ov, out->q, out->t, out->pose_score, s);
  }
  if (rc) return rc;
  hipLaunchKernelGGL(k_align_finish, dim3((batch->n_pairs + 255) / 256), dim3(256), 0, s, batch->n_pairs, B,
                     out->q, out->t, out->pose_score, out->best_score, out->best_idx, out->best_tf);
  return check_launch("fss_align (finish)");
}

// Tiling checks and launch geometry of the persistent optimiser (shared by fss_align and
// fss_align_noalloc).  The outputs other than pose_score are checked here.
struct AlignPlan {
  int block;
  LdsPlan lp;
  size_t lds;
  AlignK ak;
  long grid;
  int variant;   // the inner loop launched: ap->variant, or the expanded / streamed form when the
                 // clouds do not fit the requested layout in LDS
};

int check_align_out(const FssAlignOut* out) {
  if (!out->q || !out->t || !out->best_score || !out->best_idx || !out->best_tf || !out->n_steps)
    return set_error(FSS_E_INVALID, "NULL align outputs");
  return FSS_OK;
}

int plan_align(const FssPairBatch* batch, const KScore& ks, const FssAlignParams* ap, AlignPlan& pl) {
  const int B = ap->n_restarts, lpp = ap->lanes_per_pose, rpw = ap->restarts_per_wg;
  if (B <= 0) return set_error(FSS_E_INVALID, "n_restarts must be > 0");
  if (ap->max_steps < 0) return set_error(FSS_E_INVALID, "max_steps must be >= 0");
  if (rpw < 1) return set_error(FSS_E_INVALID, "restarts_per_wg must be >= 1");
  if (lpp < 1 || lpp > 1024 || (lpp > 64 && (lpp & (lpp - 1))))
    return set_error(FSS_E_INVALID, "lanes_per_pose must be in [1, 64], or 128, 256, 512 or 1024 on the streamed form");
  if (ap->early_stop && rpw < B)
    return set_error(FSS_E_INVALID, "early stop couples all %d restarts of a pair: restarts_per_wg (%d) must be >= n_restarts", B, rpw);
  const int groups = ap->groups_per_wg > 0 ? ap->groups_per_wg : rpw;
  if (groups > rpw) return set_error(FSS_E_INVALID, "groups_per_wg (%d) exceeds restarts_per_wg (%d)", groups, rpw);
  const int ppw = ap->pairs_per_wg > 0 ? ap->pairs_per_wg : 1;
  if (ppw > 1 && rpw < B)
    return set_error(FSS_E_INVALID, "pairs_per_wg > 1 puts every restart of each pair in the workgroup: "
                                    "restarts_per_wg (%d) must be >= n_restarts (%d)", rpw, B);
  const long lanes = (long)ppw * groups * lpp;
  if ((lpp & (lpp - 1)) && lanes > 64)
    return set_error(FSS_E_INVALID, "lanes_per_pose must be a power of two unless the workgroup's "
                                    "pairs_per_wg * groups_per_wg * lanes_per_pose lanes fit one wavefront (64)");
  if (lanes > 1024)
    return set_error(FSS_E_INVALID, "%d pairs x %d lane groups of %d lanes exceed 1024 threads: lower "
                                    "pairs_per_wg or groups_per_wg", ppw, groups, lpp);
  if (ppw > 1 && lanes > 512)
    return set_error(FSS_E_INVALID, "a workgroup of %d pairs holds at most 512 lanes (%ld asked)", ppw, lanes);
  pl.block = (int)((lanes + 63) / 64) * 64;
  int var = effective_variant(ks.mode, ap->variant);
  if (var == 3 && ppw > 1)
    return set_error(FSS_E_INVALID, "pairs_per_wg > 1 is for the VALU inner loops (variant 1 or 2)");
  if (var == 3 && lpp != 64)
    return set_error(FSS_E_INVALID, "the MFMA inner loop (variant 3) runs one pose per wavefront: lanes_per_pose must be 64");
  if (ap->tile_rows < 0) return set_error(FSS_E_INVALID, "tile_rows must be >= 0");
  const size_t cap = max_lds_bytes();
  const int sls = (rpw + groups - 1) / groups > 1 ? 24 : 14;   // floats of a restart's slot
  // lane groups of several wavefronts sum through 13 floats of LDS per wavefront
  const bool xwave = lpp > 64;
  const size_t tail = ppw * ((ap->early_stop ? 2 * (size_t)rpw * sizeof(double) : 0) + (size_t)rpw * sls * sizeof(float)) +
                      (xwave ? 13 * sizeof(float) * (size_t)(pl.block / 64) : 0);
  // Several pairs: their LDS regions start 16 B apart modulo 128 B (a region of 8k + 1 float4s),
  // so the lanes of different pairs reading the same row index hit different banks.
  auto region_of = [&](const LdsPlan& lp) { return ppw > 1 ? lp.end + ((1 - lp.end) % 8 + 8) % 8 : lp.end; };
  pl.lp = plan_lds(batch, ks.mode, var == 3);
  int region = region_of(pl.lp);
  pl.lds = (size_t)ppw * region * sizeof(float4) + tail;
  if (var == 3 && pl.lds > cap) {
    // the MFMA layout takes 76 B per ref row; clouds that fit the expanded form's 20 B run it
    const LdsPlan e = plan_lds(batch, ks.mode, false);
    if ((size_t)e.end * sizeof(float4) + tail <= cap) {
      var = 2;
      pl.lp = e;
      region = e.end;
      pl.lds = (size_t)region * sizeof(float4) + tail;
    }
  }
  // Staged clouds whose LDS leaves a CU fewer than 8 wavefronts run faster streamed (measured:
  // 7000 x 300 surface + ESP rows, 517 ms staged at 2 wavefronts per CU, 191 ms streamed), so
  // the default inner loop streams them too.
  bool low_occupancy = false;
  if (tileable(ks.mode) && ap->variant == 0 && var == 2 && ppw == 1 && pl.lds <= cap) {
    const long wpg = pl.block / 64;
    low_occupancy = std::min(16L, (long)(kCuLdsBytes / pl.lds) * wpg) < 8;
  }
  if (tileable(ks.mode) && (var == 4 || pl.lds > cap || low_occupancy)) {
    // Streamed clouds: each optimiser step passes over (ref tile, fit tile) pairs staged in LDS,
    // so a pair of any size runs; the tiles are restaged only when a side has more than one.
    if (ppw > 1)
      return set_error(FSS_E_TOO_LARGE, "clouds of up to %d x %d rows do not fit LDS: the streamed path runs one "
                                        "pair per workgroup (pairs_per_wg 1)", batch->max_ref, batch->max_fit);
    if (var == 1)
      return set_error(FSS_E_TOO_LARGE, "clouds of up to %d x %d rows do not fit LDS, and the streamed path has "
                                        "the expanded form only (variant 0, 2 or 4)", batch->max_ref, batch->max_fit);
    const int t0 = ap->tile_rows > 0 ? ap->tile_rows : 1024;
    int tr = std::max(1, std::min(batch->max_ref, t0)), tf = std::max(1, std::min(batch->max_fit, t0));
    while ((size_t)plan_lds_tiled(batch, ks.mode, tr, tf).end * sizeof(float4) + tail > cap && std::max(tr, tf) > 16) {
      if (tr >= tf) tr = (tr + 1) / 2;
      else tf = (tf + 1) / 2;
    }
    var = 4;
    pl.lp = plan_lds_tiled(batch, ks.mode, tr, tf);
    region = pl.lp.end;
    pl.lds = (size_t)region * sizeof(float4) + tail;
  }
  // Lane groups of a non-power-of-two width sum through LDS (about a sixth of the shuffle tree's
  // instructions per step), unless the scratch would cost occupancy: with it a workgroup must
  // still fit 16 to a CU (4 one-wavefront workgroups per SIMD, the VGPR limit).
  const size_t red = (size_t)red_floats(ppw * groups, lpp) * sizeof(float);
  const bool red_lds = (lpp & (lpp - 1)) && var != 4 && pl.lds + red <= kCuLdsBytes / 16;
  if (red_lds) pl.lds += red;
  // lane groups of several wavefronts: the streamed form only (every lane meets every barrier)
  if (xwave && var != 4)
    return set_error(FSS_E_INVALID, "lanes_per_pose above 64 runs on the streamed form only (variant 4, or "
                                    "clouds beyond LDS in the shape modes)");
  if (pl.lds > cap) return set_error(FSS_E_TOO_LARGE, "clouds need %zu B of LDS, the device allows %zu B", pl.lds, cap);
  pl.variant = var;
  AlignK& ak = pl.ak;
  ak.B = B; ak.max_steps = ap->max_steps;
  ak.lr = ap->lr; ak.b1 = ap->beta1; ak.b2 = ap->beta2; ak.eps = ap->eps;
  {
    double b1p = 1.0, b2p = 1.0;
    for (int k = 0; k < kSchedSteps; ++k) {
      b1p *= (double)ak.b1;
      b2p *= (double)ak.b2;
      ak.sched[2 * k] = (float)((double)ak.lr / (1.0 - b1p));
      ak.sched[2 * k + 1] = (float)(1.0 / std::sqrt(1.0 - b2p));
    }
    ak.b1p_tail = b1p;
    ak.b2p_tail = b2p;
  }
  ak.es_tol = ap->es_tol; ak.es_pat = ap->es_patience;
  ak.lpp = lpp; ak.rpw = rpw; ak.wg_per_pair = (B + rpw - 1) / rpw;
  ak.groups = groups;
  ak.red_lds = red_lds;
  ak.xwave = xwave;
  ak.ppw = ppw;
  ak.region = region;
  ak.order = ap->pair_order;
  pl.grid = ppw > 1 ? ((long)batch->n_pairs + ppw - 1) / ppw : (long)batch->n_pairs * ak.wg_per_pair;
  if (pl.grid > 0x7fffffffL) return set_error(FSS_E_INVALID, "grid too large");
  return FSS_OK;
}

// The optimiser launch and (unless skip_finish) the re-scoring and best-restart finish, with the
// self-overlaps and the per-restart score buffer given explicitly.
int run_align(const FssPairBatch* batch, const KScore& ks, const FssAlignParams* ap, AlignPlan pl,
              const double* self_ov, double* pose_score, const FssAlignOut* out, hipStream_t s) {
  pl.ak.self_ov = self_ov;
  pl.ak.q = out->q; pl.ak.t = out->t; pl.ak.pose_score = pose_score; pl.ak.n_steps = out->n_steps;
  const bool es = ap->early_stop != 0;
#define FSS_ALIGN(M) \
  (es ? launch_align<M, true>(batch, ks, pl.lp, pl.ak, (int)pl.grid, pl.block, pl.lds, pl.variant, s) \
      : launch_align<M, false>(batch, ks, pl.lp, pl.ak, (int)pl.grid, pl.block, pl.lds, pl.variant, s))
  switch (ks.mode) {
    case FSS_MODE_SHAPE: FSS_ALIGN(FSS_MODE_SHAPE); break;
    case FSS_MODE_SHAPE_ESP: FSS_ALIGN(FSS_MODE_SHAPE_ESP); break;
    case FSS_MODE_PHARM: FSS_ALIGN(FSS_MODE_PHARM); break;
    case FSS_MODE_ESP_COMBO: FSS_ALIGN(FSS_MODE_ESP_COMBO); break;
    default: FSS_ALIGN(FSS_MODE_SHAPE_PHARM); break;
  }
#undef FSS_ALIGN
  int rc;
  if ((rc = check_launch("fss_align"))) return rc;
  if (ap->skip_finish) return FSS_OK;
  FssAlignOut fo = *out;
  fo.pose_score = pose_score;
  return align_finish_impl(batch, ks, ap->n_restarts, self_ov, &fo, s);
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// No-allocation alignment (fss_align_noalloc).  The self-overlaps (4 doubles per pair) and the
// per-restart re-scoring buffer (1 double per pose) live in this file-scope device array, one copy
// per GPU, instead of in buffers the caller allocates.  Host code never passes `scratch` itself to
// a kernel (on the host the name is only a stand-in): each device's copy is looked up once with
// hipGetSymbolAddress and that device address is used.  Every call on a GPU shares the array and
// overwrites what it uses (the self-overlap and re-scoring kernels write every entry they later
// read), so nothing is cleared between calls; calls that overlap in time on one GPU race.
// ------------------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[FSS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kScratchDevices] = {};

// This GPU's copy of `scratch`: resolved on the first call on the current device, with that
// device current, and kept for the life of the process.
int scratch_address(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error((int)e, "fss_align_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kScratchDevices)
    return set_error(FSS_E_INVALID, "fss_align_noalloc: device %d is beyond the %d devices it tracks", dev,
                     kScratchDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(e != hipSuccess ? (int)e : FSS_E_NO_DEVICE, "fss_align_noalloc: hipGetSymbolAddress: %s",
                       hipGetErrorString(e));
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return FSS_OK;
}

constexpr size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// ================================================================================================
// C ABI
// ================================================================================================
extern "C" {

const char* fss_version(void) { return "fss_hip 0.1.0 gfx950"; }

const char* fss_last_error(void) { return fss::last_error(); }

int fss_device_count(int* count) {
  if (!count) return set_error(FSS_E_INVALID, "count is NULL");
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) { *count = 0; return set_error((int)e, "hipGetDeviceCount: %s", hipGetErrorString(e)); }
  return FSS_OK;
}

int fss_self_overlap(const FssPairBatch* batch, const FssScoreParams* sp, double* out, void* stream) {
  KScore ks;
  int rc = make_kscore(sp, ks);
  if (rc) return rc;
  if ((rc = check_batch(batch, ks.mode))) return rc;
  if (!out) return set_error(FSS_E_INVALID, "out is NULL");
  if (batch->n_pairs == 0) return FSS_OK;
  hipStream_t s = (hipStream_t)stream;
  switch (ks.mode) {
    case FSS_MODE_SHAPE: launch_self<FSS_MODE_SHAPE>(batch, ks, out, s); break;
    case FSS_MODE_SHAPE_ESP: launch_self<FSS_MODE_SHAPE_ESP>(batch, ks, out, s); break;
    case FSS_MODE_PHARM: launch_self<FSS_MODE_PHARM>(batch, ks, out, s); break;
    case FSS_MODE_ESP_COMBO: launch_self<FSS_MODE_ESP_COMBO>(batch, ks, out, s); break;
    default: launch_self<FSS_MODE_SHAPE_PHARM>(batch, ks, out, s); break;
  }
  return check_launch("fss_self_overlap");
}

// LDS plan of the one-workgroup-per-pose kernels: the whole pair when it fits, else (shape modes)
// streamed tiles of up to 1024 rows per side.
static int plan_pose_lds(const FssPairBatch* batch, int mode, LdsPlan& lp) {
  lp = plan_lds(batch, mode);
  const size_t cap = max_lds_bytes() - 2 * (kPoseBlock / 64) * 13 * sizeof(float);
  // staged whole when it fits and leaves a CU at least 8 wavefronts
  const size_t need = (size_t)lp.end * sizeof(float4);
  if (need <= cap && (!tileable(mode) || (kCuLdsBytes / std::max<size_t>(need, 1)) * (kPoseBlock / 64) >= 8))
    return FSS_OK;
  if (tileable(mode)) {
    lp = plan_lds_tiled(batch, mode, std::max(1, std::min(batch->max_ref, 1024)), std::max(1, std::min(batch->max_fit, 1024)));
    if ((size_t)lp.end * sizeof(float4) <= cap) return FSS_OK;
  }
  if (need <= cap) {
    lp = plan_lds(batch, mode);
    return FSS_OK;
  }
  return set_error(FSS_E_TOO_LARGE, "clouds need %zu B of LDS, the device allows %zu B",
                   (size_t)lp.end * sizeof(float4), cap);
}

int fss_overlap_grad(const FssPairBatch* batch, const FssScoreParams* sp, int32_t n_poses,
                     const int32_t* pose_pair, const float* q, const float* t, int32_t need_grad,
                     float* V, float* gR, float* gT, float* gQ, void* stream) {
  KScore ks;
  int rc = make_kscore(sp, ks);
  if (rc) return rc;
  if ((rc = check_batch(batch, ks.mode))) return rc;
  if (n_poses < 0 || !pose_pair || !q || !t || !V) return set_error(FSS_E_INVALID, "NULL pose arrays");
  if (need_grad && (!gR || !gT || !gQ)) return set_error(FSS_E_INVALID, "NULL gradient outputs");
  if (n_poses == 0) return FSS_OK;
  LdsPlan lp;
  if ((rc = plan_pose_lds(batch, ks.mode, lp))) return rc;
  hipStream_t s = (hipStream_t)stream;
#define FSS_POSE0(M)                                                                                      \
  do {                                                                                                    \
    if (need_grad) launch_pose<M, true, 0>(batch, ks, lp, n_poses, pose_pair, q, t, nullptr, 1.f, V, gR, gT, gQ, nullptr, nullptr, s); \
    else launch_pose<M, false, 0>(batch, ks, lp, n_poses, pose_pair, q, t, nullptr, 1.f, V, gR, gT, gQ, nullptr, nullptr, s);         \
  } while (0)
  switch (ks.mode) {
    case FSS_MODE_SHAPE: FSS_POSE0(FSS_MODE_SHAPE); break;
    case FSS_MODE_SHAPE_ESP: FSS_POSE0(FSS_MODE_SHAPE_ESP); break;
    case FSS_MODE_PHARM: FSS_POSE0(FSS_MODE_PHARM); break;
    case FSS_MODE_ESP_COMBO: FSS_POSE0(FSS_MODE_ESP_COMBO); break;
    default: FSS_POSE0(FSS_MODE_SHAPE_PHARM); break;
  }
#undef FSS_POSE0
  return check_launch("fss_overlap_grad");
}

int fss_pose_loss_grad(const FssPairBatch* batch, const FssScoreParams* sp, int32_t n_poses,
                       const int32_t* pose_pair, int32_t n_restarts, const double* self_ov,
                       const float* q, const float* t, float* loss, float* grad, void* stream) {
  KScore ks;
  int rc = make_kscore(sp, ks);
  if (rc) return rc;
  if ((rc = check_batch(batch, ks.mode))) return rc;
  if (n_poses < 0 || !pose_pair || !q || !t || !loss || !grad || !self_ov)
    return set_error(FSS_E_INVALID, "NULL pose arrays");
  if (n_restarts <= 0) return set_error(FSS_E_INVALID, "n_restarts must be > 0");
  if (n_poses == 0) return FSS_OK;
  LdsPlan lp;
  if ((rc = plan_pose_lds(batch, ks.mode, lp))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const float invB = 1.0f / (float)n_restarts;
  switch (ks.mode) {
    case FSS_MODE_SHAPE: launch_pose<FSS_MODE_SHAPE, true, 1>(batch, ks, lp, n_poses, pose_pair, q, t, self_ov, invB, nullptr, nullptr, nullptr, nullptr, loss, grad, s); break;
    case FSS_MODE_SHAPE_ESP: launch_pose<FSS_MODE_SHAPE_ESP, true, 1>(batch, ks, lp, n_poses, pose_pair, q, t, self_ov, invB, nullptr, nullptr, nullptr, nullptr, loss, grad, s); break;
    case FSS_MODE_PHARM: launch_pose<FSS_MODE_PHARM, true, 1>(batch, ks, lp, n_poses, pose_pair, q, t, self_ov, invB, nullptr, nullptr, nullptr, nullptr, loss, grad, s); break;
    case FSS_MODE_ESP_COMBO: launch_pose<FSS_MODE_ESP_COMBO, true, 1>(batch, ks, lp, n_poses, pose_pair, q, t, self_ov, invB, nullptr, nullptr, nullptr, nullptr, loss, grad, s); break;
    default: launch_pose<FSS_MODE_SHAPE_PHARM, true, 1>(batch, ks, lp, n_poses, pose_pair, q, t, self_ov, invB, nullptr, nullptr, nullptr, nullptr, loss, grad, s); break;
  }
  return check_launch("fss_pose_loss_grad");
}

int fss_adam_step(int32_t n, float* q, float* t, const float* grad, float* m, float* v, int32_t step,
                  float lr, float beta1, float beta2, float eps, const int32_t* pair_active,
                  int32_t n_restarts, void* stream) {
  if (n < 0 || !q || !t || !grad || !m || !v) return set_error(FSS_E_INVALID, "NULL Adam arrays");
  if (step < 1) return set_error(FSS_E_INVALID, "Adam step count is 1-based");
  if (pair_active && n_restarts <= 0) return set_error(FSS_E_INVALID, "n_restarts must be > 0");
  if (n == 0) return FSS_OK;
  const int tot = n * 7;
  hipLaunchKernelGGL(k_adam, dim3((tot + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, q, t, grad, m, v,
                     step, lr, beta1, beta2, eps, pair_active, pair_active ? n_restarts : 1);
  return check_launch("fss_adam_step");
}

int fss_adam_qt(int32_t n, float* q, float* t, const float* dQ, const float* dT, float* m_q, float* v_q,
                float* m_t, float* v_t, double lr, double beta1, double beta2, double eps,
                int32_t tangent_proj, void* stream) {
  if (n < 0 || (n > 0 && (!q || !t || !dQ || !dT || !m_q || !v_q || !m_t || !v_t)))
    return set_error(FSS_E_INVALID, "NULL Adam arrays");
  if (n == 0) return FSS_OK;
  // 1 - beta in double, as the reference forms the moment weights from Python floats
  hipLaunchKernelGGL(k_adam_qt, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, q, t, dQ, dT,
                     m_q, v_q, m_t, v_t, (float)lr, (float)beta1, (float)(1.0 - beta1), (float)beta2,
                     (float)(1.0 - beta2), (float)eps, (int)(tangent_proj != 0));
  return check_launch("fss_adam_qt");
}

int fss_early_stop_update(int32_t n_pairs, int32_t n_restarts, const float* loss, double* last,
                          int32_t* counter, int32_t* active, int32_t* n_steps, int32_t step, float tol,
                          int32_t patience, void* stream) {
  if (n_pairs < 0 || n_restarts <= 0 || !loss || !last || !counter || !active || !n_steps)
    return set_error(FSS_E_INVALID, "bad early-stop arguments");
  if (n_pairs == 0) return FSS_OK;
  hipLaunchKernelGGL(k_es_update, dim3((n_pairs + 255) / 256), dim3(256), 0, (hipStream_t)stream, n_pairs,
                     n_restarts, loss, last, counter, active, n_steps, step, (double)tol, patience);
  return check_launch("fss_early_stop_update");
}

int fss_pose_score(const FssPairBatch* batch, const FssScoreParams* sp, int32_t n_poses,
                   const int32_t* pose_pair, int32_t n_restarts, const double* self_ov, const float* q,
                   const float* t, double* score, void* stream) {
  KScore ks;
  int rc = make_kscore(sp, ks);
  if (rc) return rc;
  if ((rc = check_batch(batch, ks.mode))) return rc;
  if (n_poses < 0 || !self_ov || !q || !t || !score) return set_error(FSS_E_INVALID, "NULL pose arrays");
  if (!pose_pair && n_restarts <= 0) return set_error(FSS_E_INVALID, "n_restarts must be > 0 without pose_pair");
  if (n_poses == 0) return FSS_OK;
  return pose_score_impl(batch, ks, n_poses, pose_pair, n_restarts, self_ov, q, t, score, (hipStream_t)stream);
}

int fss_align(const FssPairBatch* batch, const FssScoreParams* sp, const FssAlignParams* ap,
              const double* self_ov, FssAlignOut* out, void* stream) {
  KScore ks;
  int rc = make_kscore(sp, ks);
  if (rc) return rc;
  if ((rc = check_batch(batch, ks.mode))) return rc;
  if (!ap || !out || !self_ov) return set_error(FSS_E_INVALID, "NULL align arguments");
  if (!out->pose_score) return set_error(FSS_E_INVALID, "NULL align outputs");
  if ((rc = check_align_out(out))) return rc;
  AlignPlan pl;
  if ((rc = plan_align(batch, ks, ap, pl))) return rc;
  if (batch->n_pairs == 0) return FSS_OK;
  return run_align(batch, ks, ap, pl, self_ov, out->pose_score, out, (hipStream_t)stream);
}

int fss_align_plan(const FssPairBatch* batch, const FssScoreParams* sp, const FssAlignParams* ap,
                   FssAlignPlanInfo* info) {
  KScore ks;
  int rc = make_kscore(sp, ks);
  if (rc) return rc;
  if ((rc = check_batch(batch, ks.mode))) return rc;
  if (!ap || !info) return set_error(FSS_E_INVALID, "NULL align arguments");
  AlignPlan pl;
  if ((rc = plan_align(batch, ks, ap, pl))) return rc;
  info->variant = pl.variant;
  info->tile_ref = pl.lp.tile_ref;
  info->tile_fit = pl.lp.tile_fit;
  info->block = pl.block;
  info->grid = pl.grid;
  info->lds_bytes = (int64_t)pl.lds;
  return FSS_OK;
}

int fss_align_noalloc(const FssPairBatch* batch, const FssScoreParams* sp, const FssAlignParams* ap,
                      FssAlignOut* out, void* stream) {
  KScore ks;
  int rc = make_kscore(sp, ks);
  if (rc) return rc;
  if ((rc = check_batch(batch, ks.mode))) return rc;
  if (!ap || !out) return set_error(FSS_E_INVALID, "NULL align arguments");
  if (ap->skip_finish)
    return set_error(FSS_E_INVALID, "fss_align_noalloc always finishes (its self-overlaps stay in the scratch "
                                    "buffer): skip_finish must be 0");
  if ((rc = check_align_out(out))) return rc;
  AlignPlan pl;
  if ((rc = plan_align(batch, ks, ap, pl))) return rc;
  const size_t so_bytes = round256((size_t)batch->n_pairs * 4 * sizeof(double));
  const size_t need = so_bytes + (size_t)batch->n_pairs * (size_t)ap->n_restarts * sizeof(double);
  if (need > (size_t)FSS_NOALLOC_SCRATCH_BYTES)
    return set_error(FSS_E_TOO_LARGE,
                     "fss_align_noalloc: %d pairs x %d restarts need %zu B of scratch; the library holds %zu B "
                     "per GPU (at most about %zu poses): split the batch or use fss_align",
                     batch->n_pairs, ap->n_restarts, need, (size_t)FSS_NOALLOC_SCRATCH_BYTES,
                     (size_t)FSS_NOALLOC_SCRATCH_BYTES / sizeof(double));
  if (batch->n_pairs == 0) return FSS_OK;
  unsigned char* base = nullptr;
  if ((rc = scratch_address(&base))) return rc;
  double* so = reinterpret_cast<double*>(base);
  double* ps = reinterpret_cast<double*>(base + so_bytes);
  hipStream_t s = (hipStream_t)stream;
  switch (ks.mode) {
    case FSS_MODE_SHAPE: launch_self<FSS_MODE_SHAPE>(batch, ks, so, s); break;
    case FSS_MODE_SHAPE_ESP: launch_self<FSS_MODE_SHAPE_ESP>(batch, ks, so, s); break;
    case FSS_MODE_PHARM: launch_self<FSS_MODE_PHARM>(batch, ks, so, s); break;
    case FSS_MODE_ESP_COMBO: launch_self<FSS_MODE_ESP_COMBO>(batch, ks, so, s); break;
    default: launch_self<FSS_MODE_SHAPE_PHARM>(batch, ks, so, s); break;
  }
  if ((rc = check_launch("fss_align_noalloc (self-overlaps)"))) return rc;
  return run_align(batch, ks, ap, pl, so, ps, out, s);
}

int fss_align_finish(const FssPairBatch* batch, const FssScoreParams* sp, int32_t n_restarts,
                     const double* self_ov, FssAlignOut* out, void* stream) {
  KScore ks;
  int rc = make_kscore(sp, ks);
  if (rc) return rc;
  if ((rc = check_batch(batch, ks.mode))) return rc;
  if (!out || !self_ov || !out->q || !out->t || !out->pose_score || !out->best_score || !out->best_idx || !out->best_tf)
    return set_error(FSS_E_INVALID, "NULL align outputs");
  if (n_restarts <= 0) return set_error(FSS_E_INVALID, "n_restarts must be > 0");
  if (batch->n_pairs == 0) return FSS_OK;
  return align_finish_impl(batch, ks, n_restarts, self_ov, out, (hipStream_t)stream);
}

int fss_seed_poses(const FssPairBatch* batch, int32_t flags, int32_t n_restarts,
                   const float* quat_table, float* q, float* t, void* stream) {
  if (!batch || !quat_table || !q || !t) return set_error(FSS_E_INVALID, "NULL seed arguments");
  if (n_restarts <= 0) return set_error(FSS_E_INVALID, "n_restarts must be > 0");
  if ((flags & FSS_SEED_PHARM) && (flags & FSS_SEED_SURF))
    return set_error(FSS_E_INVALID, "FSS_SEED_PHARM and FSS_SEED_SURF are exclusive");
  // Surface seeds: the kernel reads the ShaEP channel's surface rows through the shape slots.
  FssPairBatch b = *batch;
  if (flags & FSS_SEED_SURF) {
    b.ref_pts = batch->ref_srf; b.ref_rng = batch->ref_srng;
    b.fit_pts = batch->fit_srf; b.fit_rng = batch->fit_srng;
  }
  const bool use_pharm = flags & FSS_SEED_PHARM;
  if (use_pharm ? (!b.ref_prng || !b.fit_prng || !b.ref_anc || !b.fit_anc)
                : (!b.ref_rng || !b.fit_rng || !b.ref_pts || !b.fit_pts))
    return set_error(FSS_E_INVALID, "cloud arrays are NULL");
  if (b.n_pairs <= 0) return FSS_OK;
  hipLaunchKernelGGL(k_seed_poses, dim3(b.n_pairs), dim3(64), 0, (hipStream_t)stream, b,
                     (int)(flags & (FSS_SEED_PHARM | FSS_SEED_REF)), n_restarts, quat_table, q, t);
  return check_launch("fss_seed_poses");
}

int fss_score_fwd(const FssPairBatch* batch, const FssScoreParams* sp, float* S, float* ov, void* stream) {
  if (!batch) return set_error(FSS_E_INVALID, "pair batch is NULL");
  return score_fwd_t<float>(batch, sp, batch->ref_pts, batch->fit_pts, S, ov, (hipStream_t)stream);
}

int fss_score_bwd(const FssPairBatch* batch, const FssScoreParams* sp, const float* ov, const float* gS,
                  float* d_fit, float* d_ref, void* stream) {
  if (!batch) return set_error(FSS_E_INVALID, "pair batch is NULL");
  return score_bwd_t<float>(batch, sp, batch->ref_pts, batch->fit_pts, ov, gS, d_fit, d_ref, (hipStream_t)stream);
}

int fss_score_fwd_f64(const FssPairBatch* batch, const FssScoreParams* sp, const double* ref_pts,
                      const double* fit_pts, double* S, double* ov, void* stream) {
  return score_fwd_t<double>(batch, sp, ref_pts, fit_pts, S, ov, (hipStream_t)stream);
}

int fss_score_bwd_f64(const FssPairBatch* batch, const FssScoreParams* sp, const double* ref_pts,
                      const double* fit_pts, const double* ov, const double* gS, double* d_fit,
                      double* d_ref, void* stream) {
  return score_bwd_t<double>(batch, sp, ref_pts, fit_pts, ov, gS, d_fit, d_ref, (hipStream_t)stream);
}

}  // extern "C"
