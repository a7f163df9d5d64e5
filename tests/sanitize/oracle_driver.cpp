// Drives the oracle's C++ CPU baseline (oracle/cpu_baseline.cpp + oracle/cpu_kernels.cpp) for the
// address / undefined-behaviour sanitizer build of tests/sanitize/build.py (test infrastructure):
// self-overlaps and multi-restart alignments of small uneven batches in the shape, shape + ESP,
// pharmacophore and shape + pharmacophore modes, on 1 and on 4 threads.  Checks that every score
// is finite and in [0, 1], that the best restart is the argmax of the per-restart scores, that the
// step counts are within the limit, and that the thread count does not change any result.  Exit
// status 0 on success.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <random>
#include <vector>

#include "../../include/fss.h"

extern "C" {
int fss_cpu_self_overlap(const FssPairBatch* pb, const FssScoreParams* sp, double* out, int n_threads);
int fss_cpu_align(const FssPairBatch* pb, const FssScoreParams* sp, const FssAlignParams* ap,
                  const double* self_ov, float* q, float* t, double* pose_score, double* best_score,
                  int32_t* best_idx, int32_t* n_steps, int n_threads);
}

namespace {

int g_failures = 0;

void fail(const char* what, int mode, long a = 0) {
  fprintf(stderr, "FAIL mode %d: %s (%ld)\n", mode, what, a);
  ++g_failures;
}

struct Side {
  std::vector<float> pts, anc, vec;
  std::vector<int32_t> rng, prng, typ;
};

Side make_side(std::mt19937& g, int P, int lo, int hi, int plo, int phi) {
  std::uniform_real_distribution<float> u(-2.5f, 2.5f), w(-0.3f, 0.3f);
  std::uniform_int_distribution<int> n(lo, hi), np(plo, phi), ty(0, FSS_N_PTYPES - 2);
  Side s;
  for (int p = 0; p < P; ++p) {
    const int k = n(g), m = np(g);
    s.rng.push_back((int32_t)s.pts.size() / 4);
    s.rng.push_back(k);
    for (int i = 0; i < k; ++i) { s.pts.push_back(u(g)); s.pts.push_back(u(g)); s.pts.push_back(u(g)); s.pts.push_back(w(g)); }
    s.prng.push_back((int32_t)s.typ.size());
    s.prng.push_back(m);
    for (int i = 0; i < m; ++i) {
      for (int c = 0; c < 3; ++c) s.anc.push_back(u(g));
      s.anc.push_back(0.f);
      float v[3] = {u(g), u(g), u(g)};
      for (int c = 0; c < 3; ++c) s.vec.push_back(v[c]);
      s.vec.push_back(0.f);
      s.typ.push_back(ty(g));
    }
  }
  return s;
}

struct Run {
  std::vector<double> so, pose, best;
  std::vector<int32_t> idx, steps;
  std::vector<float> q, t;
};

Run align(const FssPairBatch& pb, const FssScoreParams& sp, const FssAlignParams& ap, const std::vector<float>& q0,
          int threads) {
  const int P = pb.n_pairs, B = ap.n_restarts;
  Run r;
  r.so.assign(4 * (size_t)P, 0.0);
  r.pose.assign((size_t)P * B, 0.0);
  r.best.assign(P, 0.0);
  r.idx.assign(P, -1);
  r.steps.assign(P, -1);
  r.q = q0;
  r.t.assign(3 * (size_t)P * B, 0.f);
  fss_cpu_self_overlap(&pb, &sp, r.so.data(), threads);
  fss_cpu_align(&pb, &sp, &ap, r.so.data(), r.q.data(), r.t.data(), r.pose.data(), r.best.data(), r.idx.data(),
                r.steps.data(), threads);
  return r;
}

void run_mode(int mode, unsigned seed) {
  const int P = 7, B = 6;
  std::mt19937 g(seed);
  const Side ref = make_side(g, P, 1, 40, 1, 9), fit = make_side(g, P, 1, 35, 1, 7);
  FssPairBatch pb{};
  pb.n_pairs = P;
  pb.ref_pts = ref.pts.data(); pb.fit_pts = fit.pts.data(); pb.ref_rng = ref.rng.data(); pb.fit_rng = fit.rng.data();
  pb.ref_anc = ref.anc.data(); pb.ref_vec = ref.vec.data(); pb.ref_typ = ref.typ.data(); pb.ref_prng = ref.prng.data();
  pb.fit_anc = fit.anc.data(); pb.fit_vec = fit.vec.data(); pb.fit_typ = fit.typ.data(); pb.fit_prng = fit.prng.data();
  pb.max_ref = pb.max_fit = 40;
  pb.max_ref_ph = pb.max_fit_ph = 9;
  FssScoreParams sp{};
  sp.mode = mode;
  sp.similarity = FSS_SIM_TANIMOTO;
  sp.alpha = 0.81;
  sp.lam = 0.09;
  sp.color_weight = 0.5;
  sp.sigma = 0.5;
  const int32_t cats[FSS_N_PTYPES] = {1, 1, 2, 0, 1, 0, 0, 0, 3};
  for (int i = 0; i < FSS_N_PTYPES; ++i) {
    sp.ph_alpha[i] = 1.0;
    sp.ph_k[i] = pow(M_PI / 2.0, 1.5);
    sp.ph_cat[i] = cats[i];
  }
  FssAlignParams ap{};
  ap.n_restarts = B; ap.max_steps = 40; ap.lr = 0.1f; ap.beta1 = 0.9f; ap.beta2 = 0.999f; ap.eps = 1e-8f;
  ap.early_stop = 1; ap.es_tol = 1e-5f; ap.es_patience = 10;
  std::normal_distribution<float> nd;
  std::vector<float> q0(4 * (size_t)P * B);
  for (size_t i = 0; i < q0.size(); i += 4) {
    float n2 = 0.f;
    for (int k = 0; k < 4; ++k) { q0[i + k] = nd(g); n2 += q0[i + k] * q0[i + k]; }
    for (int k = 0; k < 4; ++k) q0[i + k] /= sqrtf(n2);
  }
  const Run a = align(pb, sp, ap, q0, 1), b = align(pb, sp, ap, q0, 4);
  for (int p = 0; p < P; ++p) {
    double mx = -1.0;
    int arg = -1;
    for (int r = 0; r < B; ++r) {
      const double s = a.pose[(size_t)p * B + r];
      if (!(s >= -1e-12 && s <= 1.0 + 1e-9)) fail("score outside [0, 1]", mode, p);
      if (s > mx) { mx = s; arg = r; }
    }
    if (a.best[p] != mx || a.idx[p] != arg) fail("best restart is not the argmax", mode, p);
    if (a.steps[p] < 1 || a.steps[p] > ap.max_steps) fail("step count", mode, a.steps[p]);
  }
  if (a.pose != b.pose || a.q != b.q || a.t != b.t || a.steps != b.steps || a.so != b.so)
    fail("1 and 4 threads differ", mode);
  printf("mode %d: %d pairs x %d restarts, best[0] %.6f after %d steps\n", mode, P, B, a.best[0], a.steps[0]);
}

}  // namespace

int main() {
  run_mode(FSS_MODE_SHAPE, 21);
  run_mode(FSS_MODE_SHAPE_ESP, 22);
  run_mode(FSS_MODE_PHARM, 23);
  run_mode(FSS_MODE_SHAPE_PHARM, 24);
  if (g_failures) {
    fprintf(stderr, "%d failures\n", g_failures);
    return 1;
  }
  printf("oracle driver: OK\n");
  return 0;
}
