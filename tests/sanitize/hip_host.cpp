// Host stand-in of the HIP runtime calls and device entry points that the screening service
// (fast_shepherd_score_amd/csrc/fss_screen.cpp) makes, so that the service's host code -- the
// submit FIFO, admission by row pools, launch planning, the worker and poll threads, completion
// and shutdown -- builds with g++ and runs under the address, undefined-behaviour and thread
// sanitizers on a machine without a GPU.  Test infrastructure only: never linked into
// libfss_hip.so.
//
// "Device" memory is host memory, every copy is a memcpy, streams are synchronous.  The device
// entry points stand in as follows:
//   fss_seed_poses / fss_self_overlap: fill their outputs with constants;
//   fss_align: each pair's best score is standin::score of its rows, its step count
//     standin::steps of that score, its transform the score plus the element index;
//   fss::topk_merge: the merge's contract -- the top K of the previous hits and the new pair
//     scores by (score descending, library index ascending), NaN last.
#include <hip/hip_runtime_api.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "../../fast_shepherd_score_amd/csrc/fss_topk.h"
#include "../../include/fss.h"
#include "standin.h"

namespace {
std::atomic<int> g_streams{0};
}

extern "C" {

hipError_t hipMalloc(void** p, size_t n) {
  *p = n ? calloc(n, 1) : nullptr;
  return (n && !*p) ? hipErrorOutOfMemory : hipSuccess;
}
hipError_t hipFree(void* p) {
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind) {
  if (n) memmove(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(dst, src, n, k);
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(new int(g_streams.fetch_add(1)));
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete reinterpret_cast<int*>(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "stand-in HIP error"; }

int fss_seed_poses(const FssPairBatch* pb, int32_t, int32_t B, const float* table, float* q, float* t, void*) {
  for (long i = 0; i < (long)pb->n_pairs * B; ++i) {
    memcpy(q + 4 * i, table + 4 * (i % B), sizeof(float) * 4);
    t[3 * i] = t[3 * i + 1] = t[3 * i + 2] = 0.f;
  }
  return FSS_OK;
}

int fss_self_overlap(const FssPairBatch* pb, const FssScoreParams*, double* out, void*) {
  for (long i = 0; i < 4L * pb->n_pairs; ++i) out[i] = 1.0;
  return FSS_OK;
}

int fss_align(const FssPairBatch* pb, const FssScoreParams*, const FssAlignParams*, const double*,
              FssAlignOut* out, void*) {
  for (int p = 0; p < pb->n_pairs; ++p) {
    const bool sh = pb->ref_rng && pb->ref_pts, ph = pb->ref_prng && pb->ref_anc && pb->max_ref_ph > 0;
    const int nq = sh ? pb->ref_rng[2 * p + 1] : 0, ne = sh ? pb->fit_rng[2 * p + 1] : 0;
    const int nqp = ph ? pb->ref_prng[2 * p + 1] : 0, nep = ph ? pb->fit_prng[2 * p + 1] : 0;
    const double s = standin::score(sh ? pb->ref_pts + 4L * pb->ref_rng[2 * p] : nullptr, nq,
                                    sh ? pb->fit_pts + 4L * pb->fit_rng[2 * p] : nullptr, ne,
                                    ph ? pb->ref_anc + 4L * pb->ref_prng[2 * p] : nullptr,
                                    ph ? pb->ref_typ + pb->ref_prng[2 * p] : nullptr, nqp,
                                    ph ? pb->fit_anc + 4L * pb->fit_prng[2 * p] : nullptr,
                                    ph ? pb->fit_typ + pb->fit_prng[2 * p] : nullptr, nep);
    out->best_score[p] = s;
    out->best_idx[p] = 0;
    out->n_steps[p] = standin::steps(s);
    for (int k = 0; k < 16; ++k) out->best_tf[16 * p + k] = (float)s + (float)k;
  }
  return FSS_OK;
}

}  // extern "C"

namespace fss {

hipError_t topk_merge(int K, const MergeJob* jobs, int nj, const double* new_s, const float* new_tf,
                      hipStream_t) {
  struct Cand { double s; int64_t i; const float* tf; };
  for (int j = 0; j < nj; ++j) {
    const MergeJob& jb = jobs[j];
    std::vector<Cand> c;
    for (int r = 0; r < jb.prev_n; ++r) c.push_back({jb.prev_s[r], jb.prev_i[r], jb.prev_tf + 16 * r});
    for (int r = 0; r < jb.k; ++r) c.push_back({new_s[jb.off + r], jb.base + r, new_tf + 16 * (jb.off + r)});
    std::stable_sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) {
      if (a.s != a.s || b.s != b.s) return b.s != b.s && a.s == a.s;   // NaN last
      return a.s > b.s || (a.s == b.s && a.i < b.i);
    });
    const int need = std::min<int>(K, (int)c.size());
    for (int r = 0; r < need; ++r) {
      jb.out_s[r] = c[r].s;
      jb.out_i[r] = c[r].i;
      memcpy(jb.out_tf + 16 * r, c[r].tf, sizeof(float) * 16);
    }
  }
  return hipSuccess;
}

}  // namespace fss
