// CPU baseline of the ShEPhERD alignment hot path: C++17, a persistent std::thread pool.
//
// TEST INFRASTRUCTURE ONLY (the oracle package): timed beside the GPU by bench.py's cpu_baseline
// leg and checked against the FP64 oracle by tests; never part of the product path.
//
// It restates the reference's numba CPU backend -- the SoA pair-term kernels
// accel/kernels/cpu_soa.py:27-106 (in cpu_kernels.cpp: FP32 terms, vector exp, FP64
// accumulation), the pharmacophore kernel cpu.py:500-582 (FP64), and the fine loop of
// cpu_fused.py:179-217 (every step: the kernel over all poses in parallel, then the tail) --
// driving the optimiser semantics of the HIP path (alignment/_torch_analytical.py:23-478:
// torch-style Adam on raw (q, t), gradient / B, per-pair early stop on the restart-mean loss,
// FP64 re-scoring of the final poses, argmax).
// Parallelism: numba's prange over poses.  A pool of std::threads (no OpenMP runtime, so the
// library can share a process with torch's) takes the still-active poses of every pair from an
// atomic counter each step; the per-pair early-stop bookkeeping runs between steps.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../include/fss.h"

extern "C" void fss_cpu_pose_sums(const float* ax, const float* ay, const float* az, const float* ac, int n,
                                  const float* fit, int m, const float* R, const float* t, float half_alpha,
                                  float inv_lam, int esp, double* out);

namespace {

// A fixed pool: run(n, fn) calls fn(i) for i in [0, n) on the pool's threads and the caller,
// returning when every call has finished.
class Pool {
 public:
  explicit Pool(int n_threads) {
    for (int k = 1; k < n_threads; ++k) threads_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& th : threads_) th.join();
  }
  void run(int n, const std::function<void(int)>& fn) {
    if (n <= 0) return;
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      busy_ = (int)threads_.size();
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work() {
    for (int i = next_.fetch_add(1); i < n_; i = next_.fetch_add(1)) (*fn_)(i);
  }
  void loop() {
    long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      work();
      std::lock_guard<std::mutex> lk(mu_);
      if (--busy_ == 0) done_cv_.notify_one();
    }
  }
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* fn_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, busy_ = 0;
  long gen_ = 0;
  bool stop_ = false;
};

struct Cloud { const float* p; int n; };
struct PhCloud { const float* anc; const float* vec; const int32_t* typ; int n; };

void rotation(const float* q, double qh[4], double& qn, double R[9]) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  qn = sqrt(w * w + x * x + y * y + z * z);
  qh[0] = w / qn; qh[1] = x / qn; qh[2] = y / qn; qh[3] = z / qn;
  const double a = qh[0], b = qh[1], c = qh[2], d = qh[3];
  R[0] = 1 - 2 * (c * c + d * d); R[1] = 2 * (b * c - d * a);     R[2] = 2 * (b * d + c * a);
  R[3] = 2 * (b * c + d * a);     R[4] = 1 - 2 * (b * b + d * d); R[5] = 2 * (c * d - b * a);
  R[6] = 2 * (b * d - c * a);     R[7] = 2 * (c * d + b * a);     R[8] = 1 - 2 * (b * b + c * c);
}

void project(const double G[9], const double qh[4], double g[4]) {
  const double w = qh[0], x = qh[1], y = qh[2], z = qh[3];
  g[0] = -2 * z * G[1] + 2 * y * G[2] + 2 * z * G[3] - 2 * x * G[5] - 2 * y * G[6] + 2 * x * G[7];
  g[1] = 2 * y * G[1] + 2 * z * G[2] + 2 * y * G[3] - 4 * x * G[4] - 2 * w * G[5] + 2 * z * G[6] + 2 * w * G[7] - 4 * x * G[8];
  g[2] = -4 * y * G[0] + 2 * x * G[1] + 2 * w * G[2] + 2 * x * G[3] + 2 * z * G[5] - 2 * w * G[6] + 2 * z * G[7] - 4 * y * G[8];
  g[3] = -4 * z * G[0] - 2 * w * G[1] + 2 * x * G[2] + 2 * w * G[3] - 4 * z * G[4] + 2 * y * G[5] + 2 * x * G[6] + 2 * y * G[7];
}

// A shape cloud as structure-of-arrays (x, y, z, ESP), the layout the vector kernel reads.
struct Soa {
  std::vector<float> x, y, z, c;
  int n = 0;
};

Soa to_soa(const Cloud& A) {
  Soa s;
  s.n = A.n;
  s.x.resize(A.n); s.y.resize(A.n); s.z.resize(A.n); s.c.resize(A.n);
  for (int i = 0; i < A.n; ++i) {
    s.x[i] = A.p[4 * i]; s.y[i] = A.p[4 * i + 1]; s.z[i] = A.p[4 * i + 2]; s.c[i] = A.p[4 * i + 3];
  }
  return s;
}

// O, dO/dt, dO/dR of the shape channel: the vector kernel's FP64 sums, scaled by K and alpha K.
void shape_grad(const Soa& A, const Cloud& Bc, const float R[9], const float t[3], float alpha, bool esp,
                float inv_lam, double& O, double gt[3], double G[9]) {
  double s[13];
  fss_cpu_pose_sums(A.x.data(), A.y.data(), A.z.data(), A.c.data(), A.n, Bc.p, Bc.n, R, t, 0.5f * alpha,
                    inv_lam, esp ? 1 : 0, s);
  const double K = pow(M_PI / (2.0 * alpha), 1.5), aK = K * alpha;
  O = s[0] * K;
  for (int i = 0; i < 3; ++i) gt[i] = s[1 + i] * aK;
  for (int i = 0; i < 9; ++i) G[i] = s[4 + i] * aK;
}

void pharm_grad(const PhCloud& A, const PhCloud& Bc, const float R[9], const float t[3],
                const FssScoreParams& sp, double& O, double gt[3], double G[9]) {
  O = 0; for (int i = 0; i < 3; ++i) gt[i] = 0; for (int i = 0; i < 9; ++i) G[i] = 0;
  for (int b = 0; b < Bc.n; ++b) {
    const int tb = Bc.typ[b];
    if (tb < 0 || tb >= FSS_N_PTYPES || sp.ph_cat[tb] == 3) continue;
    const int cat = sp.ph_cat[tb];
    const double al = sp.ph_alpha[tb], K = sp.ph_k[tb];
    const float* p = Bc.anc + 4 * b;
    const float* u0 = Bc.vec + 4 * b;
    double un = sqrt((double)u0[0] * u0[0] + (double)u0[1] * u0[1] + (double)u0[2] * u0[2]);
    un = un > 1e-12 ? un : 1e-12;
    const double ux0 = u0[0] / un, uy0 = u0[1] / un, uz0 = u0[2] / un;
    const double bx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + t[0];
    const double by = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + t[1];
    const double bz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + t[2];
    const double ux = R[0] * ux0 + R[1] * uy0 + R[2] * uz0;
    const double uy = R[3] * ux0 + R[4] * uy0 + R[5] * uz0;
    const double uz = R[6] * ux0 + R[7] * uy0 + R[8] * uz0;
    double fx = 0, fy = 0, fz = 0, hx = 0, hy = 0, hz = 0;
    for (int a = 0; a < A.n; ++a) {
      if (A.typ[a] != tb) continue;
      const float* r = A.anc + 4 * a;
      const float* rv = A.vec + 4 * a;
      double rn = sqrt((double)rv[0] * rv[0] + (double)rv[1] * rv[1] + (double)rv[2] * rv[2]);
      rn = rn > 1e-12 ? rn : 1e-12;
      const double vx = rv[0] / rn, vy = rv[1] / rn, vz = rv[2] / rn;
      const double dx = r[0] - bx, dy = r[1] - by, dz = r[2] - bz;
      const double e = exp(-0.5 * al * (dx * dx + dy * dy + dz * dz));
      const double D = ux * vx + uy * vy + uz * vz;
      double w = 1, c = 0;
      if (cat == 1) { w = (fmin(fmax(D, 0.0), 1.0) + 2) / 3; c = (D > 0 && D < 1) ? 1 : 0; }
      else if (cat == 2) { w = (fabs(D) + 2) / 3; c = D > 0 ? 1 : (D < 0 ? -1 : 0); }
      else if (cat == 5) w = 0;
      O += K * w * e;
      const double fw = al * K * w * e;
      fx += fw * dx; fy += fw * dy; fz += fw * dz;
      const double hc = K / 3.0 * e * c;
      hx += hc * vx; hy += hc * vy; hz += hc * vz;
      if (cat >= 4) {
        // extended points anchor + unit vector (ag/_torch.py:333-357): the fit point moves as
        // R (p + u0) + t, so its force adds to both the anchor and the vector outer products
        const double ex = (r[0] + vx) - (bx + ux), ey = (r[1] + vy) - (by + uy), ez = (r[2] + vz) - (bz + uz);
        const double e2 = exp(-0.5 * al * (ex * ex + ey * ey + ez * ez));
        O += K * e2;
        const double fe = al * K * e2;
        fx += fe * ex; fy += fe * ey; fz += fe * ez;
        hx += fe * ex; hy += fe * ey; hz += fe * ez;
      }
    }
    gt[0] += fx; gt[1] += fy; gt[2] += fz;
    G[0] += fx * p[0] + hx * ux0; G[1] += fx * p[1] + hx * uy0; G[2] += fx * p[2] + hx * uz0;
    G[3] += fy * p[0] + hy * ux0; G[4] += fy * p[1] + hy * uy0; G[5] += fy * p[2] + hy * uz0;
    G[6] += fz * p[0] + hz * ux0; G[7] += fz * p[1] + hz * uy0; G[8] += fz * p[2] + hz * uz0;
  }
}

double shape_value64(const Cloud& A, const Cloud& Bc, const double R[9], const double t[3],
                     double alpha, bool esp, double lam) {
  double v = 0;
  for (int b = 0; b < Bc.n; ++b) {
    const float* p = Bc.p + 4 * b;
    const double bx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + t[0];
    const double by = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + t[1];
    const double bz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + t[2];
    for (int a = 0; a < A.n; ++a) {
      const float* r = A.p + 4 * a;
      const double dx = r[0] - bx, dy = r[1] - by, dz = r[2] - bz;
      double arg = -0.5 * alpha * (dx * dx + dy * dy + dz * dz);
      if (esp) { const double dc = (double)r[3] - (double)p[3]; arg -= dc * dc / lam; }
      v += exp(arg);
    }
  }
  return v * pow(M_PI / (2.0 * alpha), 1.5);
}

double pharm_value64(const PhCloud& A, const PhCloud& Bc, const double R[9], const double t[3],
                     const FssScoreParams& sp) {
  float Rf[9], tf[3];
  for (int i = 0; i < 9; ++i) Rf[i] = (float)R[i];
  for (int i = 0; i < 3; ++i) tf[i] = (float)t[i];
  double O, gt[3], G[9];
  pharm_grad(A, Bc, Rf, tf, sp, O, gt, G);
  return O;
}

struct Sim { double S, ws, wp; };

Sim similarity(const FssScoreParams& sp, double Os, double Op, const double so[4], bool sh, bool ph) {
  Sim r{0, 0, 0};
  const double cs = sp.mode == FSS_MODE_SHAPE_PHARM ? 1.0 - sp.color_weight : 1.0;
  const double cp = sp.mode == FSS_MODE_SHAPE_PHARM ? sp.color_weight : 1.0;
  if (sh) {
    const double U = so[0] + so[1], D = U - Os;
    r.S += cs * Os / D; r.ws = cs * U / (D * D);
  }
  if (ph) {
    if (sp.similarity == FSS_SIM_TANIMOTO) {
      const double U = so[2] + so[3], D = U - Op;
      r.S += cp * Op / D; r.wp = cp * U / (D * D);
    } else {
      const double D = sp.sigma * so[2] + (1.0 - sp.sigma) * so[3];
      r.S += cp * fmin(Op / D, 1.0); r.wp = Op < D ? cp / D : 0.0;
    }
  }
  return r;
}

}  // namespace

extern "C" {

// Self-overlaps (n_pairs, 4) in FP64.
int fss_cpu_self_overlap(const FssPairBatch* pb, const FssScoreParams* sp, double* out, int n_threads) {
  const bool sh = sp->mode != FSS_MODE_PHARM, esp = sp->mode == FSS_MODE_SHAPE_ESP;
  const bool ph = sp->mode == FSS_MODE_PHARM || sp->mode == FSS_MODE_SHAPE_PHARM;
  const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, z[3] = {0, 0, 0};
  Pool pool(n_threads);
  pool.run(2 * pb->n_pairs, [&](int k) {
    const int p = k >> 1, side = k & 1;
    double vs = 0, vp = 0;
    if (sh) {
      const int32_t* rg = side ? pb->fit_rng : pb->ref_rng;
      Cloud c{(side ? pb->fit_pts : pb->ref_pts) + 4 * (long)rg[2 * p], rg[2 * p + 1]};
      vs = shape_value64(c, c, I, z, sp->alpha, esp, sp->lam);
    }
    if (ph) {
      const int32_t* rg = side ? pb->fit_prng : pb->ref_prng;
      const long o = rg[2 * p];
      PhCloud c{(side ? pb->fit_anc : pb->ref_anc) + 4 * o, (side ? pb->fit_vec : pb->ref_vec) + 4 * o,
                (side ? pb->fit_typ : pb->ref_typ) + o, rg[2 * p + 1]};
      vp = pharm_value64(c, c, I, z, *sp);
    }
    out[4 * p + side] = vs;
    out[4 * p + 2 + side] = vp;
  });
  return 0;
}

// Full multi-restart alignment of every pair (host pointers; the layout of fss_align).  Every
// step runs the loss + gradient + Adam update of all still-active poses in parallel, then the
// per-pair early-stop test; the final poses are re-scored in FP64, also in parallel.
int fss_cpu_align(const FssPairBatch* pb, const FssScoreParams* sp, const FssAlignParams* ap,
                  const double* self_ov, float* q, float* t, double* pose_score, double* best_score,
                  int32_t* best_idx, int32_t* n_steps, int n_threads) {
  const bool sh = sp->mode != FSS_MODE_PHARM, esp = sp->mode == FSS_MODE_SHAPE_ESP;
  const bool ph = sp->mode == FSS_MODE_PHARM || sp->mode == FSS_MODE_SHAPE_PHARM;
  const int P = pb->n_pairs, B = ap->n_restarts;
  const float inv_lam = esp ? (float)(1.0 / sp->lam) : 0.f;
  Pool pool(n_threads);
  std::vector<Soa> ref(sh ? P : 0);
  std::vector<Cloud> fit(sh ? P : 0);
  std::vector<PhCloud> pref(ph ? P : 0), pfit(ph ? P : 0);
  pool.run(P, [&](int p) {
    if (sh) {
      ref[p] = to_soa(Cloud{pb->ref_pts + 4 * (long)pb->ref_rng[2 * p], pb->ref_rng[2 * p + 1]});
      fit[p] = Cloud{pb->fit_pts + 4 * (long)pb->fit_rng[2 * p], pb->fit_rng[2 * p + 1]};
    }
    if (ph) {
      const long ro = pb->ref_prng[2 * p], fo = pb->fit_prng[2 * p];
      pref[p] = PhCloud{pb->ref_anc + 4 * ro, pb->ref_vec + 4 * ro, pb->ref_typ + ro, pb->ref_prng[2 * p + 1]};
      pfit[p] = PhCloud{pb->fit_anc + 4 * fo, pb->fit_vec + 4 * fo, pb->fit_typ + fo, pb->fit_prng[2 * p + 1]};
    }
  });
  const long n_poses = (long)P * B;
  std::vector<float> mom(14 * n_poses, 0.f);          // Adam m (7) and v (7) per pose
  std::vector<double> loss(n_poses, 0.0);
  std::vector<double> last(P, 1.0);
  std::vector<int> counter(P, 0);
  std::vector<int32_t> live;                          // active pairs
  for (int p = 0; p < P; ++p) { live.push_back(p); n_steps[p] = ap->max_steps; }
  double b1p = 1, b2p = 1;
  for (int step = 1; step <= ap->max_steps && !live.empty(); ++step) {
    b1p *= ap->beta1; b2p *= ap->beta2;
    const float step_size = (float)(ap->lr / (1.0 - b1p));
    const float ibc2s = (float)(1.0 / sqrt(1.0 - b2p));
    pool.run((int)live.size() * B, [&](int k) {
      const int p = live[k / B], r = k % B;
      const long pose = (long)p * B + r;
      float* qp = q + 4 * pose;
      float* tp = t + 3 * pose;
      double qh[4], qn, Rd[9];
      rotation(qp, qh, qn, Rd);
      float R[9];
      for (int i = 0; i < 9; ++i) R[i] = (float)Rd[i];
      double Os = 0, Op = 0, gts[3] = {0, 0, 0}, Gs[9] = {0}, gtp[3] = {0, 0, 0}, Gp[9] = {0};
      if (sh) shape_grad(ref[p], fit[p], R, tp, (float)sp->alpha, esp, inv_lam, Os, gts, Gs);
      if (ph) pharm_grad(pref[p], pfit[p], R, tp, *sp, Op, gtp, Gp);
      const Sim s = similarity(*sp, Os, Op, self_ov + 4 * p, sh, ph);
      loss[pose] = 1.0 - s.S;
      double lR[9], lT[3];
      for (int i = 0; i < 9; ++i) lR[i] = -(s.ws * Gs[i] + s.wp * Gp[i]);
      for (int i = 0; i < 3; ++i) lT[i] = -(s.ws * gts[i] + s.wp * gtp[i]);
      double gq[4];
      project(lR, qh, gq);
      const double dot = qh[0] * gq[0] + qh[1] * gq[1] + qh[2] * gq[2] + qh[3] * gq[3];
      float g[7];
      for (int i = 0; i < 4; ++i) g[i] = (float)((gq[i] - qh[i] * dot) / qn / B);
      for (int i = 0; i < 3; ++i) g[4 + i] = (float)(lT[i] / B);
      float* mm = &mom[14 * pose];
      float* vv = mm + 7;
      for (int i = 0; i < 7; ++i) {
        float* par = i < 4 ? qp + i : tp + (i - 4);
        mm[i] = mm[i] + (1.f - ap->beta1) * (g[i] - mm[i]);
        vv[i] = vv[i] * ap->beta2 + (1.f - ap->beta2) * g[i] * g[i];
        *par = *par - step_size * (mm[i] * (1.f / std::fma(sqrtf(vv[i]), ibc2s, ap->eps)));
      }
    });
    if (!ap->early_stop) continue;
    std::vector<int32_t> still;
    for (int p : live) {
      double L = 0;
      for (int r = 0; r < B; ++r) L += loss[(long)p * B + r];
      L /= B;
      counter[p] = fabs(L - last[p]) > ap->es_tol ? 0 : counter[p] + 1;
      last[p] = L;
      if (counter[p] > ap->es_patience) n_steps[p] = step;
      else still.push_back(p);
    }
    live.swap(still);
  }
  pool.run((int)n_poses, [&](int k) {
    const int p = k / B;
    double qh[4], qn, Rd[9];
    rotation(q + 4 * (long)k, qh, qn, Rd);
    const double td[3] = {t[3 * (long)k], t[3 * (long)k + 1], t[3 * (long)k + 2]};
    Cloud A = sh ? Cloud{pb->ref_pts + 4 * (long)pb->ref_rng[2 * p], pb->ref_rng[2 * p + 1]} : Cloud{nullptr, 0};
    const double vs = sh ? shape_value64(A, fit[p], Rd, td, sp->alpha, esp, sp->lam) : 0.0;
    const double vp = ph ? pharm_value64(pref[p], pfit[p], Rd, td, *sp) : 0.0;
    pose_score[k] = similarity(*sp, vs, vp, self_ov + 4 * p, sh, ph).S;
  });
  for (int p = 0; p < P; ++p) {
    int bi = 0;
    for (int r = 1; r < B; ++r)
      if (pose_score[(long)p * B + r] > pose_score[(long)p * B + bi]) bi = r;
    best_score[p] = pose_score[(long)p * B + bi];
    best_idx[p] = bi;
  }
  return 0;
}

}  // extern "C"
