// Screening service of libfss_hip.so: a conformer library resident in HBM and a worker thread that
// aligns queued queries against all of it with continuous batching (the C ABI is in fss.h).
//
// Structure (restating the reference's screen / _TopK contract, shepherd_score/screen.py:898-1192,
// over a resident library instead of streamed shards):
//   * fss_screen_submit copies the query into a host FIFO and wakes the worker;
//   * the worker admits queries into free slots (query rows copied into the slot's device arena),
//     then repeatedly fills one launch of up to launch_pairs pairs from the admitted queries,
//     oldest first (fss_plan_launch: the tail of one query and the head of the next share a
//     launch), runs fss_seed_poses -> fss_self_overlap -> fss_align on it and merges
//     every query's new pair scores into its top-k (k_topk_merge, double-buffered per slot);
//   * a query whose library range is exhausted has its top-k copied to the host, and pollers
//     waiting on its ticket are woken.
// Pair results depend only on the pair (fixed tiling, per-pair seeds and early stop), and the
// merge orders by (score descending, library index ascending), so a query's hits are the same
// whatever it shares launches with.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/fss.h"
#include "fss_common.h"
#include "fss_topk.h"

using fss::MergeJob;
using fss::set_error;

struct FssLibrary {
  int device = 0, channels = 0;
  int64_t cap_conf = 0, cap_pts = 0, cap_ph = 0;
  int64_t n = 0, n_pts = 0, n_ph = 0;
  int32_t max_pts = 0, max_ph = 0;
  float* pts = nullptr;        // (cap_pts, 4)
  int32_t* rng = nullptr;      // (cap_conf, 2)
  float* anc = nullptr;        // (cap_ph, 4)
  float* vec = nullptr;        // (cap_ph, 4)
  int32_t* typ = nullptr;      // (cap_ph)
  int32_t* prng = nullptr;     // (cap_conf, 2)
  // ShaEP channel
  int64_t cap_srf = 0, cap_atm = 0, n_srf = 0, n_atm = 0;
  int32_t max_srf = 0, max_atm = 0;
  float* srf = nullptr;        // (cap_srf, 4): surface point, ESP
  int32_t* srng = nullptr;     // (cap_conf, 2)
  float* atm = nullptr;        // (cap_atm, 4): atom with hydrogens, partial charge
  float* rad = nullptr;        // (cap_atm)
  int32_t* arng = nullptr;     // (cap_conf, 2)
  std::vector<int32_t> h_rng, h_prng, h_srng, h_arng;   // host copies of the range tables
  std::mutex mu;
  int services = 0;
};

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0, clk::time_point t1) {
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// Switches the calling thread to `dev` for the lifetime of the guard.
struct DeviceGuard {
  int prev = -1;
  hipError_t err = hipSuccess;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    err = hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

int hip_fail(hipError_t e, const char* what) {
  return set_error((int)e, "%s: %s", what, hipGetErrorString(e));
}

template <typename T>
hipError_t dalloc(T** p, int64_t count) {
  *p = nullptr;
  if (count <= 0) return hipSuccess;
  return hipMalloc((void**)p, (size_t)count * sizeof(T));
}

void free_library(FssLibrary* lib) {
  void* ps[] = {lib->pts, lib->rng, lib->anc, lib->vec, lib->typ, lib->prng,
                lib->srf, lib->srng, lib->atm, lib->rad, lib->arng};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete lib;
}

// ------------------------------------------------------------------------------------------------
// Query memory.  An admitted query's rows live in four device pools (shape rows; pharmacophore
// anchors, vectors and types; ShaEP surface rows; ShaEP atoms and radii), each a run of rows that
// a first-fit allocator hands out in contiguous ranges and merges back on release.  A query is
// admitted only when every pool has a range for its rows, so admission is by the memory it needs.
// ------------------------------------------------------------------------------------------------
class RowPool {
 public:
  void reset(int64_t rows) {
    free_.clear();
    cap_ = rows;
    if (rows > 0) free_[0] = rows;
  }
  int64_t capacity() const { return cap_; }
  // First free range of n rows (n == 0 always fits at row 0).
  bool alloc(int64_t n, int64_t* first) {
    if (n <= 0) { *first = 0; return true; }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *first = it->first;
      const int64_t rest = it->second - n, at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(int64_t first, int64_t n) {
    if (n <= 0) return;
    auto it = free_.emplace(first, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  // True when [first, first + n) lies in the pool and overlaps no free range.
  bool allocated(int64_t first, int64_t n) const {
    if (n <= 0) return n == 0;
    if (first < 0 || first + n > cap_) return false;
    auto it = free_.upper_bound(first);
    if (it != free_.end() && it->first < first + n) return false;
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second > first) return false;
    }
    return true;
  }
  void stats(int64_t* free_rows, int64_t* largest) const {
    int64_t f = 0, l = 0;
    for (const auto& kv : free_) { f += kv.second; l = std::max(l, kv.second); }
    if (free_rows) *free_rows = f;
    if (largest) *largest = l;
  }

 private:
  std::map<int64_t, int64_t> free_;   // first row -> length of each free range
  int64_t cap_ = 0;
};

// Admission, oldest first: query k of n needs rows[k * np + c] rows of pool c; it is admitted
// while a slot is free and every pool has a range for it.  The first query that does not fit
// ends the round (no later query overtakes it).  first[k * np + c] receives the ranges; returns
// the number admitted.
int admit_fifo(RowPool* const* pools, int np, int n, const int64_t* rows, int free_slots, int64_t* first) {
  int k = 0;
  for (; k < n && k < free_slots; ++k) {
    int c = 0;
    for (; c < np; ++c)
      if (!pools[c]->alloc(rows[(size_t)k * np + c], &first[(size_t)k * np + c])) break;
    if (c < np) {
      for (int d = 0; d < c; ++d) pools[d]->release(first[(size_t)k * np + d], rows[(size_t)k * np + d]);
      break;
    }
  }
  return k;
}

enum QueryPool { kQPts, kQPharm, kQSurf, kQAtoms, kQPools };
constexpr int64_t kQRowBytes[kQPools] = {16, 36, 16, 20};   // bytes of one row in each pool

struct QueryIn {
  int64_t ticket = 0;
  std::vector<float> pts, anc, vec, srf, atm, rad;
  std::vector<int32_t> typ;
  int32_t np = 0, nph = 0, ns = 0, na = 0;
  bool want_all = false;
  bool mem_wait = false;               // has waited for pool room with a slot free
  clk::time_point t_submit;
};

// The channels a mode reads, as fss_screen_create sizes them.
// shape_rows: the shape term reads the shape channel; srf: the query and library surface rows
// are read (the ShaEP score, or the surface shape modes with shape_from_surface).
struct Channels {
  bool shape, pharm, se, shape_rows, srf;
};

Channels channels_of(int mode, int shape_from_surface) {
  Channels c;
  c.shape = mode != FSS_MODE_PHARM;
  c.pharm = mode == FSS_MODE_PHARM || mode == FSS_MODE_SHAPE_PHARM;
  c.se = mode == FSS_MODE_ESP_COMBO;
  const bool surf_shape = shape_from_surface &&
      (c.se || mode == FSS_MODE_SHAPE || mode == FSS_MODE_SHAPE_ESP);
  c.shape_rows = c.shape && !surf_shape;
  c.srf = c.se || surf_shape;
  return c;
}

struct Slot {
  bool used = false;
  QueryIn q;
  int64_t first[kQPools] = {0, 0, 0, 0};   // the query's row ranges in the pools
  int64_t rows[kQPools] = {0, 0, 0, 0};
  int64_t next = 0, lib_n = 0;
  int32_t filled = 0, buf = 0;
  std::vector<double> all;             // every entry's best score, when the query asked
  clk::time_point t_start;
};

struct Result {
  int rc = FSS_OK;
  std::string err;
  std::vector<double> score;
  std::vector<int64_t> index;
  std::vector<float> tf;
  std::vector<double> all;
  double queue_ms = 0, latency_ms = 0;
};

// Range tables of one launch, each (P, 2) = [first row, count] per pair.
enum RngTable { kRef, kFit, kRefPh, kFitPh, kRefSrf, kFitSrf, kRefAtm, kFitAtm, kRngTables };

static_assert(kQPools == FSS_QPOOLS && kQPts == FSS_QPOOL_POINTS && kQPharm == FSS_QPOOL_PHARM &&
                  kQSurf == FSS_QPOOL_SURFACE && kQAtoms == FSS_QPOOL_ATOMS,
              "pool numbering of fss.h");

}  // namespace

struct FssRowPool {
  RowPool pool;
};

struct FssScreen {
  FssLibrary* lib = nullptr;
  FssScoreParams sp{};
  FssScreenParams prm{};
  int B = 0, K = 0, P = 0, A = 0;
  int need_shape = 0, need_pharm = 0;
  int need_se = 0;                     // FSS_MODE_ESP_COMBO: the ShaEP channel
  int need_srf = 0;                    // surface rows (ShaEP, or the surface shape modes)
  int shape_rows = 0;                  // the shape term reads the shape channel's rows (else the surface rows)
  hipStream_t stream = nullptr;
  // device arena
  float *q = nullptr, *t = nullptr, *table = nullptr, *best_tf = nullptr;
  double *pose_score = nullptr, *best_score = nullptr, *self_ov = nullptr;
  int32_t *best_idx = nullptr, *n_steps = nullptr;
  int32_t *d_rng = nullptr;            // kRngTables x (P, 2), see RngTable
  int32_t* h_rng = nullptr;            // pinned host staging of d_rng
  float *qpts = nullptr, *qanc = nullptr, *qvec = nullptr;   // query pools (rows of QueryPool)
  int32_t* qtyp = nullptr;
  float *qsrf = nullptr, *qatm = nullptr, *qrad = nullptr;
  RowPool pools[kQPools];              // guarded by mu
  int64_t pool_rows[kQPools] = {0, 0, 0, 0};
  std::atomic<int64_t> admitted_max{0}, memory_waits{0};
  double* tk_s = nullptr;              // (A, 2, K)
  int64_t* tk_i = nullptr;             // (A, 2, K)
  float* tk_tf = nullptr;              // (A, 2, K, 16)
  MergeJob* d_jobs = nullptr;
  MergeJob* h_jobs = nullptr;          // pinned
  int32_t* h_steps = nullptr;          // pinned: the last launch's per-pair step counts
  int32_t *d_order = nullptr, *h_order = nullptr;   // launch order by size (pairs_per_wg > 1)
  // scheduler state
  std::vector<Slot> slots;
  std::deque<QueryIn> pending;
  std::unordered_set<int64_t> inflight;
  std::unordered_map<int64_t, Result> results;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int64_t next_ticket = 1;
  std::atomic<int64_t> launches{0}, pairs{0}, pair_steps{0};
  // sum over aligned pairs of executed steps x restarts x row pairs (worker thread writes, under mu)
  double shape_terms = 0, pharm_terms = 0, esp_terms = 0;
  std::thread worker;
};

namespace {

void free_arena(FssScreen* s) {
  void* dev_ptrs[] = {s->q, s->t, s->table, s->best_tf, s->pose_score, s->best_score, s->self_ov,
                      s->best_idx, s->n_steps, s->d_rng, s->qpts, s->qanc, s->qvec, s->qtyp,
                      s->qsrf, s->qatm, s->qrad, s->tk_s, s->tk_i, s->tk_tf, s->d_jobs, s->d_order};
  for (void* p : dev_ptrs)
    if (p) (void)hipFree(p);
  if (s->h_rng) (void)hipHostFree(s->h_rng);
  if (s->h_jobs) (void)hipHostFree(s->h_jobs);
  if (s->h_steps) (void)hipHostFree(s->h_steps);
  if (s->h_order) (void)hipHostFree(s->h_order);
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

void finish(FssScreen* s, int64_t ticket, Result&& r) {
  std::lock_guard<std::mutex> lk(s->mu);
  s->inflight.erase(ticket);
  s->results[ticket] = std::move(r);
  s->cv_done.notify_all();
}

// Frees slot i: its pool ranges go back to the pools and its query is dropped; returns the
// query's ticket.
int64_t release_slot(FssScreen* s, int i) {
  Slot& sl = s->slots[i];
  const int64_t ticket = sl.q.ticket;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (int c = 0; c < kQPools; ++c) {
      s->pools[c].release(sl.first[c], sl.rows[c]);
      sl.first[c] = sl.rows[c] = 0;
    }
  }
  sl.used = false;
  sl.q = QueryIn();
  return ticket;
}

// Copies the query rows into the slot's pool ranges (stream-ordered; the slot keeps the host
// vectors alive until the next synchronisation).
int admit(FssScreen* s, int slot) {
  Slot& sl = s->slots[slot];
  const int64_t* f = sl.first;
  hipError_t e = hipSuccess;
  if (s->shape_rows && sl.q.np)
    e = hipMemcpyAsync(s->qpts + 4 * f[kQPts], sl.q.pts.data(), sizeof(float) * 4 * sl.q.np,
                       hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess && s->need_pharm && sl.q.nph) {
    e = hipMemcpyAsync(s->qanc + 4 * f[kQPharm], sl.q.anc.data(), sizeof(float) * 4 * sl.q.nph,
                       hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s->qvec + 4 * f[kQPharm], sl.q.vec.data(), sizeof(float) * 4 * sl.q.nph,
                         hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s->qtyp + f[kQPharm], sl.q.typ.data(), sizeof(int32_t) * sl.q.nph,
                         hipMemcpyHostToDevice, s->stream);
  }
  if (e == hipSuccess && s->need_srf)
    e = hipMemcpyAsync(s->qsrf + 4 * f[kQSurf], sl.q.srf.data(), sizeof(float) * 4 * sl.q.ns,
                       hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess && s->need_se) {
    e = hipMemcpyAsync(s->qatm + 4 * f[kQAtoms], sl.q.atm.data(), sizeof(float) * 4 * sl.q.na,
                       hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s->qrad + f[kQAtoms], sl.q.rad.data(), sizeof(float) * sl.q.na,
                         hipMemcpyHostToDevice, s->stream);
  }
  if (e != hipSuccess) return hip_fail(e, "query upload");
  {
    std::lock_guard<std::mutex> lk(s->lib->mu);
    sl.lib_n = s->lib->n;
  }
  sl.next = 0;
  sl.filled = 0;
  sl.buf = 0;
  if (sl.q.want_all) sl.all.assign((size_t)sl.lib_n, std::numeric_limits<double>::quiet_NaN());
  sl.t_start = clk::now();
  return FSS_OK;
}

// Hands every slot of `act` whose library range is exhausted its result (top-k copied to the
// host) and frees the slot.  A query admitted against an empty library finishes here with no
// launch and no hits.
int complete_finished(FssScreen* s, const std::vector<int>& act, clk::time_point now) {
  for (int i : act) {
    Slot& sl = s->slots[i];
    if (!sl.used || sl.next < sl.lib_n) continue;
    Result r;
    const size_t cur = ((size_t)i * 2 + sl.buf) * s->K;
    r.score.resize(sl.filled); r.index.resize(sl.filled); r.tf.resize(16 * (size_t)sl.filled);
    if (sl.filled) {
      hipError_t e = hipMemcpy(r.score.data(), s->tk_s + cur, sizeof(double) * sl.filled, hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(r.index.data(), s->tk_i + cur, sizeof(int64_t) * sl.filled, hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(r.tf.data(), s->tk_tf + 16 * cur, sizeof(float) * 16 * sl.filled, hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_fail(e, "hit download");
    }
    r.queue_ms = ms_since(sl.q.t_submit, sl.t_start);
    r.latency_ms = ms_since(sl.q.t_submit, now);
    r.all.swap(sl.all);
    finish(s, release_slot(s, i), std::move(r));
  }
  return FSS_OK;
}

// One launch over the admitted slots `act`.
int launch(FssScreen* s, const std::vector<int>& act) {
  const int na = (int)act.size();
  std::vector<int64_t> rem(na), take(na);
  for (int i = 0; i < na; ++i) rem[i] = s->slots[act[i]].lib_n - s->slots[act[i]].next;
  fss_plan_launch(na, rem.data(), s->P, take.data());
  int32_t* tab[kRngTables];
  for (int k = 0; k < kRngTables; ++k) tab[k] = s->h_rng + 2 * (size_t)s->P * k;
  // pair p of table k <- (first row, count); a channel the mode does not use gets empty ranges
  auto set = [&](int k, int p, int32_t first, int32_t count) { tab[k][2 * p] = first; tab[k][2 * p + 1] = count; };
  int p = 0, nj = 0, max_np = 0, max_nph = 0, max_ns = 0, max_na = 0;
  int32_t lib_max_pts = 0, lib_max_ph = 0, lib_max_srf = 0, lib_max_atm = 0;
  {
    std::lock_guard<std::mutex> lk(s->lib->mu);
    const FssLibrary* L = s->lib;
    lib_max_pts = L->max_pts; lib_max_ph = L->max_ph; lib_max_srf = L->max_srf; lib_max_atm = L->max_atm;
    for (int i = 0; i < na; ++i) {
      if (take[i] == 0) continue;
      Slot& sl = s->slots[act[i]];
      MergeJob& jb = s->h_jobs[nj++];
      const size_t cur = ((size_t)act[i] * 2 + sl.buf) * s->K, nxt = ((size_t)act[i] * 2 + (sl.buf ^ 1)) * s->K;
      jb.prev_s = s->tk_s + cur; jb.prev_i = s->tk_i + cur; jb.prev_tf = s->tk_tf + 16 * cur;
      jb.out_s = s->tk_s + nxt; jb.out_i = s->tk_i + nxt; jb.out_tf = s->tk_tf + 16 * nxt;
      jb.base = sl.next; jb.prev_n = sl.filled; jb.off = p; jb.k = (int32_t)take[i]; jb.pad = 0;
      max_np = std::max(max_np, sl.q.np);
      max_nph = std::max(max_nph, sl.q.nph);
      max_ns = std::max(max_ns, sl.q.ns);
      max_na = std::max(max_na, sl.q.na);
      for (int64_t j = 0; j < take[i]; ++j, ++p) {
        const size_t e = 2 * (size_t)(sl.next + j);
        const bool sh = s->shape_rows, ph = s->need_pharm, se = s->need_se, sf = s->need_srf;
        set(kRef, p, (int32_t)sl.first[kQPts], sh ? sl.q.np : 0);
        set(kFit, p, sh ? L->h_rng[e] : 0, sh ? L->h_rng[e + 1] : 0);
        set(kRefPh, p, (int32_t)sl.first[kQPharm], ph ? sl.q.nph : 0);
        set(kFitPh, p, ph ? L->h_prng[e] : 0, ph ? L->h_prng[e + 1] : 0);
        set(kRefSrf, p, (int32_t)sl.first[kQSurf], sf ? sl.q.ns : 0);
        set(kFitSrf, p, sf ? L->h_srng[e] : 0, sf ? L->h_srng[e + 1] : 0);
        set(kRefAtm, p, (int32_t)sl.first[kQAtoms], se ? sl.q.na : 0);
        set(kFitAtm, p, se ? L->h_arng[e] : 0, se ? L->h_arng[e + 1] : 0);
      }
    }
  }
  if (p == 0) return complete_finished(s, act, clk::now());
  if (s->h_order) {   // pairs by their entry's row count, largest first (stable counting sort)
    const int kc = s->need_shape ? (s->shape_rows ? kFit : kFitSrf) : kFitPh;
    int32_t mx = 0;
    for (int i = 0; i < p; ++i) mx = std::max(mx, tab[kc][2 * i + 1]);
    std::vector<int32_t> start((size_t)mx + 2, 0);
    for (int i = 0; i < p; ++i) ++start[(size_t)(mx - tab[kc][2 * i + 1]) + 1];
    for (int c = 0; c <= mx; ++c) start[c + 1] += start[c];
    for (int i = 0; i < p; ++i) s->h_order[start[(size_t)(mx - tab[kc][2 * i + 1])]++] = i;
  }
  hipError_t e = hipMemcpyAsync(s->d_rng, s->h_rng, sizeof(int32_t) * 2 * kRngTables * (size_t)s->P,
                                hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess && s->h_order)
    e = hipMemcpyAsync(s->d_order, s->h_order, sizeof(int32_t) * p, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(s->d_jobs, s->h_jobs, sizeof(MergeJob) * nj, hipMemcpyHostToDevice, s->stream);
  if (e != hipSuccess) return hip_fail(e, "launch staging");
  const int32_t* drng[kRngTables];
  for (int k = 0; k < kRngTables; ++k) drng[k] = s->d_rng + 2 * (size_t)s->P * k;
  FssPairBatch pb{};
  pb.n_pairs = p;
  if (s->shape_rows) {
    pb.ref_pts = s->qpts; pb.fit_pts = s->lib->pts;
    pb.ref_rng = drng[kRef]; pb.fit_rng = drng[kFit];
    pb.max_ref = std::max(max_np, 1); pb.max_fit = std::max(lib_max_pts, 1);
  } else if (s->need_shape) {   // the shape term on the surface rows (ShaEP, or a surface mode)
    pb.ref_pts = s->qsrf; pb.fit_pts = s->lib->srf;
    pb.ref_rng = drng[kRefSrf]; pb.fit_rng = drng[kFitSrf];
    pb.max_ref = std::max(max_ns, 1); pb.max_fit = std::max(lib_max_srf, 1);
  } else {
    pb.ref_rng = drng[kRef]; pb.fit_rng = drng[kFit];
  }
  pb.ref_anc = s->qanc; pb.ref_vec = s->qvec; pb.ref_typ = s->qtyp; pb.ref_prng = drng[kRefPh];
  pb.fit_anc = s->lib->anc; pb.fit_vec = s->lib->vec; pb.fit_typ = s->lib->typ; pb.fit_prng = drng[kFitPh];
  pb.max_ref_ph = s->need_pharm ? std::max(max_nph, 1) : 0;
  pb.max_fit_ph = s->need_pharm ? std::max(lib_max_ph, 1) : 0;
  if (s->need_se) {
    pb.ref_srf = s->qsrf; pb.fit_srf = s->lib->srf; pb.ref_srng = drng[kRefSrf]; pb.fit_srng = drng[kFitSrf];
    pb.ref_atm = s->qatm; pb.fit_atm = s->lib->atm; pb.ref_rad = s->qrad; pb.fit_rad = s->lib->rad;
    pb.ref_arng = drng[kRefAtm]; pb.fit_arng = drng[kFitAtm];
    pb.max_ref_srf = std::max(max_ns, 1); pb.max_fit_srf = std::max(lib_max_srf, 1);
    pb.max_ref_atm = std::max(max_na, 1); pb.max_fit_atm = std::max(lib_max_atm, 1);
  }
  int rc = fss_seed_poses(&pb, s->prm.seed_flags, s->B, s->table, s->q, s->t, s->stream);
  if (!rc) rc = fss_self_overlap(&pb, &s->sp, s->self_ov, s->stream);
  if (!rc) {
    FssAlignParams ap{};
    ap.n_restarts = s->B; ap.max_steps = s->prm.max_steps;
    ap.lr = s->prm.lr; ap.beta1 = s->prm.beta1; ap.beta2 = s->prm.beta2; ap.eps = s->prm.eps;
    ap.early_stop = s->prm.early_stop; ap.es_tol = s->prm.es_tol; ap.es_patience = s->prm.es_patience;
    ap.lanes_per_pose = s->prm.lanes_per_pose; ap.restarts_per_wg = s->prm.restarts_per_wg;
    ap.pair_order = s->d_order; ap.skip_finish = 0; ap.variant = s->prm.variant;
    ap.pairs_per_wg = s->prm.pairs_per_wg;
    ap.groups_per_wg = s->prm.groups_per_wg;
    FssAlignOut out{};
    out.q = s->q; out.t = s->t; out.pose_score = s->pose_score; out.best_score = s->best_score;
    out.best_idx = s->best_idx; out.best_tf = s->best_tf; out.n_steps = s->n_steps;
    rc = fss_align(&pb, &s->sp, &ap, s->self_ov, &out, s->stream);
  }
  if (rc) return rc;
  e = fss::topk_merge(s->K, s->d_jobs, nj, s->best_score, s->best_tf, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->h_steps, s->n_steps, sizeof(int32_t) * (size_t)p, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return hip_fail(e, "screening launch");
  s->launches.fetch_add(1);
  s->pairs.fetch_add(p);
  int64_t steps = 0;
  double st_shape = 0, st_pharm = 0, st_esp = 0;
  const int kShapeRef = s->shape_rows ? kRef : kRefSrf, kShapeFit = s->shape_rows ? kFit : kFitSrf;
  auto cnt = [&](int k, int i) { return (double)tab[k][2 * i + 1]; };
  for (int i = 0; i < p; ++i) {
    const double st = (double)s->h_steps[i] * s->B;
    steps += s->h_steps[i];
    if (s->need_shape) st_shape += st * cnt(kShapeRef, i) * cnt(kShapeFit, i);
    if (s->need_pharm) st_pharm += st * cnt(kRefPh, i) * cnt(kFitPh, i);
    if (s->need_se)    // each side's surface observers against the other side's atoms
      st_esp += st * (cnt(kRefSrf, i) * cnt(kFitAtm, i) + cnt(kFitSrf, i) * cnt(kRefAtm, i));
  }
  s->pair_steps.fetch_add(steps);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->shape_terms += st_shape;
    s->pharm_terms += st_pharm;
    s->esp_terms += st_esp;
  }
  const auto now = clk::now();
  for (int i = 0, off = 0; i < na; ++i) {
    if (take[i] == 0) continue;
    Slot& sl = s->slots[act[i]];
    if (sl.q.want_all) {   // the launch's pairs [off, off + take) are entries [next, next + take)
      e = hipMemcpy(sl.all.data() + sl.next, s->best_score + off, sizeof(double) * take[i],
                    hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_fail(e, "score download");
    }
    off += (int)take[i];
    sl.next += take[i];
    sl.filled = (int32_t)std::min<int64_t>(s->K, (int64_t)sl.filled + take[i]);
    sl.buf ^= 1;
  }
  return complete_finished(s, act, now);
}

void fail_active(FssScreen* s, int rc, const std::string& msg) {
  for (int i = 0; i < s->A; ++i) {
    if (!s->slots[i].used) continue;
    Result r;
    r.rc = rc;
    r.err = msg;
    finish(s, release_slot(s, i), std::move(r));
  }
}

// Admits queued queries, oldest first, into free slots while the pools have room for their rows
// (called with s->mu held).  Returns the slots admitted.
std::vector<int> admit_pending(FssScreen* s) {
  std::vector<int> free_idx, admitted;
  for (int i = 0; i < s->A; ++i)
    if (!s->slots[i].used) free_idx.push_back(i);
  const int n = (int)std::min(free_idx.size(), s->pending.size());
  if (n == 0) return admitted;
  std::vector<int64_t> rows((size_t)n * kQPools), first((size_t)n * kQPools);
  for (int k = 0; k < n; ++k) {
    const QueryIn& q = s->pending[k];
    int64_t* r = &rows[(size_t)k * kQPools];
    r[kQPts] = s->shape_rows ? q.np : 0;
    r[kQPharm] = s->need_pharm ? q.nph : 0;
    r[kQSurf] = s->need_srf ? q.ns : 0;
    r[kQAtoms] = s->need_se ? q.na : 0;
  }
  RowPool* pp[kQPools];
  for (int c = 0; c < kQPools; ++c) pp[c] = &s->pools[c];
  const int got = admit_fifo(pp, kQPools, n, rows.data(), n, first.data());
  if (got < n && !s->pending[got].mem_wait) {   // a slot is free, the pools are not
    s->pending[got].mem_wait = true;
    s->memory_waits.fetch_add(1);
  }
  for (int k = 0; k < got; ++k) {
    Slot& sl = s->slots[free_idx[k]];
    sl.q = std::move(s->pending.front());
    s->pending.pop_front();
    for (int c = 0; c < kQPools; ++c) {
      sl.first[c] = first[(size_t)k * kQPools + c];
      sl.rows[c] = rows[(size_t)k * kQPools + c];
    }
    sl.used = true;
    admitted.push_back(free_idx[k]);
  }
  const int64_t held = s->A - (int64_t)free_idx.size() + got;
  if (held > s->admitted_max.load()) s->admitted_max.store(held);
  return admitted;
}

void worker_main(FssScreen* s) {
  if (hipSetDevice(s->lib->device) != hipSuccess) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  for (;;) {
    std::vector<int> act;
    std::vector<int> admitted;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (!s->stop) admitted = admit_pending(s);
      for (int i = 0; i < s->A; ++i)
        if (s->slots[i].used) act.push_back(i);
      // oldest query first: launches are filled in submission (ticket) order
      std::sort(act.begin(), act.end(), [&](int x, int y) { return s->slots[x].q.ticket < s->slots[y].q.ticket; });
      if (act.empty()) {
        if (s->stop) break;
        // with every slot free the pools are whole and hold any accepted query, so a query is
        // left queued here only while the service stops
        s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        continue;
      }
      if (s->stop) break;
    }
    for (int i : admitted) {
      if (admit(s, i)) {
        Result r;
        r.rc = FSS_E_INVALID;
        r.err = fss::last_error();
        finish(s, release_slot(s, i), std::move(r));
        act.erase(std::remove(act.begin(), act.end(), i), act.end());
      }
    }
    if (act.empty()) continue;
    const int rc = launch(s, act);
    if (rc) {
      (void)hipStreamSynchronize(s->stream);
      fail_active(s, rc, fss::last_error());
    }
  }
  // closing: fail what is still queued or admitted
  fail_active(s, FSS_E_CLOSED, "the screening service was closed");
  std::deque<QueryIn> left;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    left.swap(s->pending);
  }
  for (QueryIn& q : left) {
    Result r;
    r.rc = FSS_E_CLOSED;
    r.err = "the screening service was closed";
    finish(s, q.ticket, std::move(r));
  }
}

}  // namespace

extern "C" {

int fss_plan_launch(int32_t n, const int64_t* remaining, int64_t capacity, int64_t* take) {
  if (n < 0 || (n > 0 && (!remaining || !take)) || capacity < 0)
    return set_error(FSS_E_INVALID, "bad launch plan arguments");
  int64_t left = capacity;
  for (int i = 0; i < n; ++i) {
    take[i] = std::min(std::max<int64_t>(remaining[i], 0), left);
    left -= take[i];
  }
  return FSS_OK;
}

int fss_plan_query_pools(const FssScoreParams* sp, const FssScreenParams* prm, int64_t rows[FSS_QPOOLS]) {
  if (!sp || !prm || !rows) return set_error(FSS_E_INVALID, "NULL pool plan arguments");
  if (prm->max_active < 1) return set_error(FSS_E_INVALID, "max_active must be >= 1");
  if (prm->query_budget_bytes < 0) return set_error(FSS_E_INVALID, "query_budget_bytes must be >= 0");
  const Channels ch = channels_of(sp->mode, prm->shape_from_surface);
  const int64_t need[kQPools] = {ch.shape_rows ? std::max(prm->max_query_points, 0) : 0,
                                 ch.pharm ? std::max(prm->max_query_pharm, 0) : 0,
                                 ch.srf ? std::max(prm->max_query_surface, 0) : 0,
                                 ch.se ? std::max(prm->max_query_atoms, 0) : 0};
  const int64_t A = prm->max_active;
  int64_t one = 0;
  for (int c = 0; c < kQPools; ++c) one += need[c] * kQRowBytes[c];
  const int64_t budget = prm->query_budget_bytes;
  if (budget == 0 || budget >= A * one) {
    for (int c = 0; c < kQPools; ++c) rows[c] = A * need[c];
  } else {
    if (budget < one)
      return set_error(FSS_E_INVALID, "query_budget_bytes %lld does not hold one query of the maximum size (%lld bytes)",
                       (long long)budget, (long long)one);
    // m maximum queries' worth of every pool, m = budget / one in [1, A)
    for (int c = 0; c < kQPools; ++c)
      rows[c] = std::min(A * need[c], std::max(need[c], (int64_t)((__int128)budget * need[c] / one)));
  }
  for (int c = 0; c < kQPools; ++c)
    if (rows[c] > INT32_MAX) return set_error(FSS_E_INVALID, "query pool %d exceeds 2^31 rows", c);
  return FSS_OK;
}

int fss_rowpool_create(int64_t rows, FssRowPool** out) {
  if (!out || rows < 0) return set_error(FSS_E_INVALID, "bad row pool arguments");
  *out = new FssRowPool();
  (*out)->pool.reset(rows);
  return FSS_OK;
}

int fss_rowpool_alloc(FssRowPool* p, int64_t n, int64_t* first) {
  if (!p || !first || n < 0) return set_error(FSS_E_INVALID, "bad row pool arguments");
  return p->pool.alloc(n, first) ? FSS_OK : FSS_PENDING;
}

int fss_rowpool_free(FssRowPool* p, int64_t first, int64_t n) {
  if (!p || !p->pool.allocated(first, n))
    return set_error(FSS_E_INVALID, "rows [%lld, %lld) are not an allocated range", (long long)first,
                     (long long)(first + n));
  p->pool.release(first, n);
  return FSS_OK;
}

int fss_rowpool_stats(const FssRowPool* p, int64_t* free_rows, int64_t* largest_free) {
  if (!p) return set_error(FSS_E_INVALID, "row pool is NULL");
  p->pool.stats(free_rows, largest_free);
  return FSS_OK;
}

int fss_rowpool_destroy(FssRowPool* p) {
  delete p;
  return FSS_OK;
}

int fss_admit_fifo(FssRowPool* const* pools, int32_t n_pools, int32_t n, const int64_t* rows,
                   int32_t free_slots, int64_t* first, int32_t* admitted) {
  if (!admitted || n < 0 || n_pools < 0 || (n > 0 && n_pools > 0 && (!pools || !rows || !first)))
    return set_error(FSS_E_INVALID, "bad admission arguments");
  std::vector<RowPool*> pp((size_t)n_pools);
  for (int c = 0; c < n_pools; ++c) {
    if (!pools[c]) return set_error(FSS_E_INVALID, "row pool %d is NULL", c);
    pp[c] = &pools[c]->pool;
  }
  for (int64_t k = 0; k < (int64_t)n * n_pools; ++k)
    if (rows[k] < 0) return set_error(FSS_E_INVALID, "negative row count");
  *admitted = admit_fifo(pp.data(), n_pools, n, rows, std::max(free_slots, 0), first);
  return FSS_OK;
}

int fss_lib_create(const FssLibSpec* spec, FssLibrary** out) {
  if (!spec || !out) return set_error(FSS_E_INVALID, "NULL library arguments");
  *out = nullptr;
  constexpr int all = FSS_LIB_SHAPE | FSS_LIB_PHARM | FSS_LIB_SHAEP | FSS_LIB_SURFACE;
  if (!(spec->channels & all) || (spec->channels & ~all))
    return set_error(FSS_E_INVALID, "channels must be a combination of FSS_LIB_SHAPE, FSS_LIB_PHARM, "
                     "FSS_LIB_SHAEP and FSS_LIB_SURFACE");
  if ((spec->channels & FSS_LIB_SHAEP) && (spec->channels & FSS_LIB_SURFACE))
    return set_error(FSS_E_INVALID, "FSS_LIB_SHAEP already holds the surface rows: drop FSS_LIB_SURFACE");
  if (spec->max_conformers < 1 || spec->max_conformers > INT32_MAX)
    return set_error(FSS_E_INVALID, "max_conformers must be in [1, 2^31)");
  if ((spec->channels & FSS_LIB_SHAPE) && (spec->max_points < 1 || spec->max_points > INT32_MAX))
    return set_error(FSS_E_INVALID, "max_points must be in [1, 2^31)");
  if ((spec->channels & FSS_LIB_PHARM) && (spec->max_pharm < 1 || spec->max_pharm > INT32_MAX))
    return set_error(FSS_E_INVALID, "max_pharm must be in [1, 2^31)");
  if ((spec->channels & FSS_LIB_SHAEP) && (spec->max_surface < 1 || spec->max_surface > INT32_MAX ||
                                           spec->max_atoms < 1 || spec->max_atoms > INT32_MAX))
    return set_error(FSS_E_INVALID, "max_surface and max_atoms must be in [1, 2^31)");
  if ((spec->channels & FSS_LIB_SURFACE) && (spec->max_surface < 1 || spec->max_surface > INT32_MAX))
    return set_error(FSS_E_INVALID, "max_surface must be in [1, 2^31)");
  DeviceGuard dg(spec->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "fss_lib_create: device");
  FssLibrary* lib = new FssLibrary();
  lib->device = spec->device;
  lib->channels = spec->channels;
  lib->cap_conf = spec->max_conformers;
  lib->cap_pts = (spec->channels & FSS_LIB_SHAPE) ? spec->max_points : 0;
  lib->cap_ph = (spec->channels & FSS_LIB_PHARM) ? spec->max_pharm : 0;
  lib->cap_srf = (spec->channels & (FSS_LIB_SHAEP | FSS_LIB_SURFACE)) ? spec->max_surface : 0;
  lib->cap_atm = (spec->channels & FSS_LIB_SHAEP) ? spec->max_atoms : 0;
  hipError_t e = hipSuccess;
  if (lib->cap_pts) {
    e = dalloc(&lib->pts, 4 * lib->cap_pts);
    if (e == hipSuccess) e = dalloc(&lib->rng, 2 * lib->cap_conf);
  }
  if (e == hipSuccess && lib->cap_ph) {
    e = dalloc(&lib->anc, 4 * lib->cap_ph);
    if (e == hipSuccess) e = dalloc(&lib->vec, 4 * lib->cap_ph);
    if (e == hipSuccess) e = dalloc(&lib->typ, lib->cap_ph);
    if (e == hipSuccess) e = dalloc(&lib->prng, 2 * lib->cap_conf);
  }
  if (e == hipSuccess && lib->cap_srf) {
    e = dalloc(&lib->srf, 4 * lib->cap_srf);
    if (e == hipSuccess) e = dalloc(&lib->srng, 2 * lib->cap_conf);
  }
  if (e == hipSuccess && lib->cap_atm) {
    e = dalloc(&lib->atm, 4 * lib->cap_atm);
    if (e == hipSuccess) e = dalloc(&lib->rad, lib->cap_atm);
    if (e == hipSuccess) e = dalloc(&lib->arng, 2 * lib->cap_conf);
  }
  if (e != hipSuccess) {
    free_library(lib);
    return hip_fail(e, "fss_lib_create: allocation");
  }
  *out = lib;
  return FSS_OK;
}

int fss_lib_upload(FssLibrary* lib, int32_t n, const float* pts, const int32_t* counts,
                   const float* anc, const float* vec, const int32_t* typ, const int32_t* pcounts,
                   int64_t* first) {
  if (lib && (lib->channels & (FSS_LIB_SHAEP | FSS_LIB_SURFACE)))
    return set_error(FSS_E_INVALID, "the library has surface rows: append with fss_lib_append");
  FssLibRows r{};
  r.pts = pts; r.counts = counts; r.anc = anc; r.vec = vec; r.typ = typ; r.pcounts = pcounts;
  return fss_lib_append(lib, n, &r, first);
}

int fss_lib_append(FssLibrary* lib, int32_t n, const FssLibRows* in, int64_t* first) {
  if (!lib || !in) return set_error(FSS_E_INVALID, "library or rows are NULL");
  if (n < 0) return set_error(FSS_E_INVALID, "n < 0");
  const bool sh = lib->channels & FSS_LIB_SHAPE, ph = lib->channels & FSS_LIB_PHARM;
  const bool se = lib->channels & FSS_LIB_SHAEP;
  const bool sf = lib->channels & (FSS_LIB_SHAEP | FSS_LIB_SURFACE);   // surface rows
  if (n > 0 && sh && (!in->pts || !in->counts)) return set_error(FSS_E_INVALID, "shape channel arrays are NULL");
  if (n > 0 && ph && (!in->anc || !in->vec || !in->typ || !in->pcounts))
    return set_error(FSS_E_INVALID, "pharmacophore channel arrays are NULL");
  if (n > 0 && sf && (!in->srf || !in->scounts)) return set_error(FSS_E_INVALID, "surface arrays are NULL");
  if (n > 0 && se && (!in->atm || !in->rad || !in->acounts))
    return set_error(FSS_E_INVALID, "ShaEP atom arrays are NULL");
  std::lock_guard<std::mutex> lk(lib->mu);
  if (first) *first = lib->n;
  if (n == 0) return FSS_OK;
  if (lib->n + n > lib->cap_conf) return set_error(FSS_E_TOO_LARGE, "library capacity %lld conformers exceeded", (long long)lib->cap_conf);
  // per channel: rows appended, running maximum count, range table
  struct Ch { bool on; const int32_t* cnt; int64_t n_rows, cap; int32_t mx; std::vector<int32_t> r; int64_t rows; };
  Ch ch[4] = {{sh, in->counts, lib->n_pts, lib->cap_pts, lib->max_pts, {}, 0},
              {ph, in->pcounts, lib->n_ph, lib->cap_ph, lib->max_ph, {}, 0},
              {sf, in->scounts, lib->n_srf, lib->cap_srf, lib->max_srf, {}, 0},
              {se, in->acounts, lib->n_atm, lib->cap_atm, lib->max_atm, {}, 0}};
  static const char* names[4] = {"points", "pharmacophores", "surface points", "atoms"};
  for (int c = 0; c < 4; ++c) {
    if (!ch[c].on) continue;
    ch[c].r.resize(2 * (size_t)n);
    int64_t o = ch[c].n_rows;
    for (int32_t i = 0; i < n; ++i) {
      const int32_t k = ch[c].cnt[i];
      if (k < 0) return set_error(FSS_E_INVALID, "negative count of %s", names[c]);
      ch[c].r[2 * i] = (int32_t)std::min<int64_t>(o, INT32_MAX);
      ch[c].r[2 * i + 1] = k;
      o += k;
      ch[c].mx = std::max(ch[c].mx, k);
    }
    ch[c].rows = o - ch[c].n_rows;
    if (o > ch[c].cap)
      return set_error(FSS_E_TOO_LARGE, "library capacity %lld %s exceeded", (long long)ch[c].cap, names[c]);
  }
  DeviceGuard dg(lib->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "fss_lib_append: device");
  hipError_t e = hipSuccess;
  auto put = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyDefault);
  };
  if (sh) {
    put(lib->pts + 4 * lib->n_pts, in->pts, sizeof(float) * 4 * ch[0].rows);
    put(lib->rng + 2 * lib->n, ch[0].r.data(), sizeof(int32_t) * 2 * n);
  }
  if (ph) {
    put(lib->anc + 4 * lib->n_ph, in->anc, sizeof(float) * 4 * ch[1].rows);
    put(lib->vec + 4 * lib->n_ph, in->vec, sizeof(float) * 4 * ch[1].rows);
    put(lib->typ + lib->n_ph, in->typ, sizeof(int32_t) * ch[1].rows);
    put(lib->prng + 2 * lib->n, ch[1].r.data(), sizeof(int32_t) * 2 * n);
  }
  if (sf) {
    put(lib->srf + 4 * lib->n_srf, in->srf, sizeof(float) * 4 * ch[2].rows);
    put(lib->srng + 2 * lib->n, ch[2].r.data(), sizeof(int32_t) * 2 * n);
  }
  if (se) {
    put(lib->atm + 4 * lib->n_atm, in->atm, sizeof(float) * 4 * ch[3].rows);
    put(lib->rad + lib->n_atm, in->rad, sizeof(float) * ch[3].rows);
    put(lib->arng + 2 * lib->n, ch[3].r.data(), sizeof(int32_t) * 2 * n);
  }
  if (e != hipSuccess) return hip_fail(e, "fss_lib_append");
  if (sh) { lib->h_rng.insert(lib->h_rng.end(), ch[0].r.begin(), ch[0].r.end()); lib->n_pts += ch[0].rows; lib->max_pts = ch[0].mx; }
  if (ph) { lib->h_prng.insert(lib->h_prng.end(), ch[1].r.begin(), ch[1].r.end()); lib->n_ph += ch[1].rows; lib->max_ph = ch[1].mx; }
  if (sf) {
    lib->h_srng.insert(lib->h_srng.end(), ch[2].r.begin(), ch[2].r.end());
    lib->n_srf += ch[2].rows; lib->max_srf = ch[2].mx;
  }
  if (se) {
    lib->h_arng.insert(lib->h_arng.end(), ch[3].r.begin(), ch[3].r.end());
    lib->n_atm += ch[3].rows; lib->max_atm = ch[3].mx;
  }
  lib->n += n;
  return FSS_OK;
}

int fss_lib_size(const FssLibrary* lib, int64_t* n) {
  if (!lib || !n) return set_error(FSS_E_INVALID, "NULL library arguments");
  std::lock_guard<std::mutex> lk(const_cast<FssLibrary*>(lib)->mu);
  *n = lib->n;
  return FSS_OK;
}

int fss_lib_destroy(FssLibrary* lib) {
  if (!lib) return FSS_OK;
  {
    std::lock_guard<std::mutex> lk(lib->mu);
    if (lib->services > 0) return set_error(FSS_E_INVALID, "the library still has %d screening service(s)", lib->services);
  }
  DeviceGuard dg(lib->device);
  free_library(lib);
  return FSS_OK;
}

int fss_screen_create(FssLibrary* lib, const FssScoreParams* sp, const FssScreenParams* prm,
                      FssScreen** out) {
  if (!lib || !sp || !prm || !out) return set_error(FSS_E_INVALID, "NULL screening arguments");
  *out = nullptr;
  if (sp->mode < 0 || sp->mode > FSS_MODE_ESP_COMBO) return set_error(FSS_E_INVALID, "unknown mode %d", sp->mode);
  const Channels ch = channels_of(sp->mode, prm->shape_from_surface);
  const bool need_shape = ch.shape, need_pharm = ch.pharm, need_se = ch.se, shape_rows = ch.shape_rows;
  if (shape_rows && !(lib->channels & FSS_LIB_SHAPE)) return set_error(FSS_E_INVALID, "the mode needs the library's shape channel");
  if (need_pharm && !(lib->channels & FSS_LIB_PHARM)) return set_error(FSS_E_INVALID, "the mode needs the library's pharmacophore channel");
  if (need_se && !(lib->channels & FSS_LIB_SHAEP)) return set_error(FSS_E_INVALID, "the mode needs the library's ShaEP channel");
  if (ch.srf && !(lib->channels & (FSS_LIB_SHAEP | FSS_LIB_SURFACE)))
    return set_error(FSS_E_INVALID, "shape_from_surface needs the library's surface rows");
  if (prm->n_restarts < 1) return set_error(FSS_E_INVALID, "n_restarts must be >= 1");
  if (prm->max_steps < 0) return set_error(FSS_E_INVALID, "max_steps must be >= 0");
  if (prm->top_k < 1 || prm->top_k > FSS_SCREEN_MAX_TOPK) return set_error(FSS_E_INVALID, "top_k must be in [1, %d]", FSS_SCREEN_MAX_TOPK);
  if (prm->launch_pairs < 1) return set_error(FSS_E_INVALID, "launch_pairs must be >= 1");
  if (prm->max_active < 1 || prm->max_active > 4096) return set_error(FSS_E_INVALID, "max_active must be in [1, 4096]");
  if (prm->launch_pairs < prm->max_active) return set_error(FSS_E_INVALID, "launch_pairs must be >= max_active");
  if (shape_rows && prm->max_query_points < 1) return set_error(FSS_E_INVALID, "max_query_points must be >= 1");
  if (ch.srf && prm->max_query_surface < 1) return set_error(FSS_E_INVALID, "max_query_surface must be >= 1");
  if (need_se && prm->max_query_atoms < 1) return set_error(FSS_E_INVALID, "max_query_atoms must be >= 1");
  if (need_pharm && prm->max_query_pharm < 1) return set_error(FSS_E_INVALID, "max_query_pharm must be >= 1");
  if (!prm->quat_table) return set_error(FSS_E_INVALID, "quat_table is NULL");
  if (prm->pairs_per_wg < 0 || prm->pairs_per_wg > 64) return set_error(FSS_E_INVALID, "pairs_per_wg must be in [0, 64]");
  if (prm->groups_per_wg < 0) return set_error(FSS_E_INVALID, "groups_per_wg must be >= 0");
  int64_t pool_rows[kQPools];
  if (int rc = fss_plan_query_pools(sp, prm, pool_rows)) return rc;
  DeviceGuard dg(lib->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "fss_screen_create: device");
  FssScreen* s = new FssScreen();
  s->lib = lib;
  s->sp = *sp;
  s->prm = *prm;
  s->B = prm->n_restarts;
  s->K = prm->top_k;
  s->P = prm->launch_pairs;
  s->A = prm->max_active;
  s->need_shape = need_shape;
  s->need_pharm = need_pharm;
  s->need_se = need_se;
  s->need_srf = ch.srf;
  s->shape_rows = shape_rows;
  if (!s->prm.lanes_per_pose) s->prm.lanes_per_pose = s->B <= 64 ? 64 / s->B : 1;   // one wavefront per pair
  if (!s->prm.restarts_per_wg) s->prm.restarts_per_wg = s->B;
  if (!shape_rows) s->prm.max_query_points = 0;
  if (!need_pharm) s->prm.max_query_pharm = 0;
  if (!ch.srf) s->prm.max_query_surface = 0;
  if (!need_se) s->prm.max_query_atoms = 0;
  s->slots.resize(s->A);
  const int64_t PB = (int64_t)s->P * s->B;
  const int64_t AK = (int64_t)s->A * 2 * s->K;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = dalloc(&s->q, 4 * PB);
  if (e == hipSuccess) e = dalloc(&s->t, 3 * PB);
  if (e == hipSuccess) e = dalloc(&s->pose_score, PB);
  if (e == hipSuccess) e = dalloc(&s->table, 4 * (int64_t)s->B);
  if (e == hipSuccess) e = dalloc(&s->best_score, s->P);
  if (e == hipSuccess) e = dalloc(&s->best_idx, s->P);
  if (e == hipSuccess) e = dalloc(&s->best_tf, 16 * (int64_t)s->P);
  if (e == hipSuccess) e = dalloc(&s->n_steps, s->P);
  if (e == hipSuccess) e = dalloc(&s->self_ov, 4 * (int64_t)s->P);
  if (e == hipSuccess) e = dalloc(&s->d_rng, 2 * kRngTables * (int64_t)s->P);
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_rng, sizeof(int32_t) * 2 * kRngTables * (size_t)s->P, hipHostMallocDefault);
  for (int c = 0; c < kQPools; ++c) {
    s->pool_rows[c] = pool_rows[c];
    s->pools[c].reset(pool_rows[c]);
  }
  if (e == hipSuccess) e = dalloc(&s->qpts, 4 * pool_rows[kQPts]);
  if (e == hipSuccess) e = dalloc(&s->qsrf, 4 * pool_rows[kQSurf]);
  if (e == hipSuccess) e = dalloc(&s->qatm, 4 * pool_rows[kQAtoms]);
  if (e == hipSuccess) e = dalloc(&s->qrad, pool_rows[kQAtoms]);
  if (e == hipSuccess) e = dalloc(&s->qanc, 4 * pool_rows[kQPharm]);
  if (e == hipSuccess) e = dalloc(&s->qvec, 4 * pool_rows[kQPharm]);
  if (e == hipSuccess) e = dalloc(&s->qtyp, pool_rows[kQPharm]);
  if (e == hipSuccess) e = dalloc(&s->tk_s, AK);
  if (e == hipSuccess) e = dalloc(&s->tk_i, AK);
  if (e == hipSuccess) e = dalloc(&s->tk_tf, 16 * AK);
  if (e == hipSuccess) e = dalloc(&s->d_jobs, s->A);
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_jobs, sizeof(MergeJob) * (size_t)s->A, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_steps, sizeof(int32_t) * (size_t)s->P, hipHostMallocDefault);
  if (e == hipSuccess && s->prm.pairs_per_wg > 1) {
    e = dalloc(&s->d_order, s->P);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_order, sizeof(int32_t) * (size_t)s->P, hipHostMallocDefault);
  }
  if (e == hipSuccess) e = hipMemcpy(s->table, prm->quat_table, sizeof(float) * 4 * s->B, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_arena(s);
    delete s;
    return hip_fail(e, "fss_screen_create: allocation");
  }
  {
    std::lock_guard<std::mutex> lk(lib->mu);
    lib->services += 1;
  }
  s->worker = std::thread(worker_main, s);
  *out = s;
  return FSS_OK;
}

int fss_screen_submit(FssScreen* s, const FssQuery* q, int64_t* ticket) {
  if (!s || !q || !ticket) return set_error(FSS_E_INVALID, "NULL submit arguments");
  if (s->shape_rows) {
    if (q->n_points < 1 || !q->pts) return set_error(FSS_E_INVALID, "the query needs shape rows");
    if (q->n_points > s->prm.max_query_points)
      return set_error(FSS_E_TOO_LARGE, "query has %d rows, the service takes %d", q->n_points, s->prm.max_query_points);
  }
  if (s->need_pharm) {
    if (q->n_pharm < 1 || !q->anc || !q->vec || !q->typ) return set_error(FSS_E_INVALID, "the query needs pharmacophores");
    if (q->n_pharm > s->prm.max_query_pharm)
      return set_error(FSS_E_TOO_LARGE, "query has %d pharmacophores, the service takes %d", q->n_pharm, s->prm.max_query_pharm);
  }
  if (s->need_srf) {
    if (q->n_surface < 1 || !q->srf) return set_error(FSS_E_INVALID, "the query needs surface rows");
    if (q->n_surface > s->prm.max_query_surface)
      return set_error(FSS_E_TOO_LARGE, "query has %d surface rows, the service takes %d", q->n_surface,
                       s->prm.max_query_surface);
  }
  if (s->need_se) {
    if (q->n_atoms < 1 || !q->atm || !q->rad) return set_error(FSS_E_INVALID, "the query needs ShaEP atom rows");
    if (q->n_atoms > s->prm.max_query_atoms)
      return set_error(FSS_E_TOO_LARGE, "query has %d atom rows, the service takes %d", q->n_atoms,
                       s->prm.max_query_atoms);
  }
  QueryIn qi;
  qi.t_submit = clk::now();
  qi.want_all = q->all_scores != 0;
  if (s->need_srf) {
    qi.ns = q->n_surface;
    qi.srf.assign(q->srf, q->srf + 4 * (size_t)q->n_surface);
  }
  if (s->need_se) {
    qi.na = q->n_atoms;
    qi.atm.assign(q->atm, q->atm + 4 * (size_t)q->n_atoms);
    qi.rad.assign(q->rad, q->rad + (size_t)q->n_atoms);
  }
  if (s->shape_rows) {
    qi.np = q->n_points;
    qi.pts.assign(q->pts, q->pts + 4 * (size_t)q->n_points);
  }
  if (s->need_pharm) {
    qi.nph = q->n_pharm;
    qi.anc.assign(q->anc, q->anc + 4 * (size_t)q->n_pharm);
    qi.vec.assign(q->vec, q->vec + 4 * (size_t)q->n_pharm);
    qi.typ.assign(q->typ, q->typ + (size_t)q->n_pharm);
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return set_error(FSS_E_CLOSED, "the screening service is closing");
    qi.ticket = s->next_ticket++;
    *ticket = qi.ticket;
    s->inflight.insert(qi.ticket);
    s->pending.push_back(std::move(qi));
  }
  s->cv_work.notify_one();
  return FSS_OK;
}

int fss_screen_poll(FssScreen* s, int64_t ticket, int32_t wait_ms, FssHits* out) {
  if (!s) return set_error(FSS_E_INVALID, "service is NULL");
  Result r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return s->results.count(ticket) > 0; };
    if (!ready()) {
      if (!s->inflight.count(ticket)) return set_error(FSS_E_INVALID, "unknown ticket %lld", (long long)ticket);
      if (wait_ms < 0) s->cv_done.wait(lk, ready);
      else if (wait_ms > 0) s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
      if (!ready()) return FSS_PENDING;
    }
    auto it = s->results.find(ticket);
    const size_t n_all = it->second.all.size();
    if (!it->second.rc && n_all && (!out || !out->all_scores || out->n_all < (int64_t)n_all))
      return set_error(FSS_E_INVALID, "the query asked for all %zu scores: pass all_scores with n_all >= %zu",
                       n_all, n_all);
    r = std::move(it->second);
    s->results.erase(it);
  }
  if (r.rc) return set_error(r.rc, "%s", r.err.c_str());
  if (!out) return FSS_OK;
  if (!r.all.empty()) memcpy(out->all_scores, r.all.data(), sizeof(double) * r.all.size());
  out->n_all = (int64_t)r.all.size();
  const int n = std::min<int>(out->capacity, (int)r.score.size());
  if (n > 0 && (!out->score || !out->index || !out->transform))
    return set_error(FSS_E_INVALID, "NULL hit arrays");
  for (int i = 0; i < n; ++i) {
    out->score[i] = r.score[i];
    out->index[i] = r.index[i];
  }
  if (n > 0) memcpy(out->transform, r.tf.data(), sizeof(float) * 16 * n);
  out->n = n;
  out->queue_ms = r.queue_ms;
  out->latency_ms = r.latency_ms;
  return FSS_OK;
}

int fss_screen_stats(FssScreen* s, int64_t* launches, int64_t* pairs, int64_t* pair_steps) {
  if (!s) return set_error(FSS_E_INVALID, "service is NULL");
  if (launches) *launches = s->launches.load();
  if (pairs) *pairs = s->pairs.load();
  if (pair_steps) *pair_steps = s->pair_steps.load();
  return FSS_OK;
}

int fss_screen_terms(FssScreen* s, double* shape_terms, double* pharm_terms, double* esp_terms) {
  if (!s) return set_error(FSS_E_INVALID, "service is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  if (shape_terms) *shape_terms = s->shape_terms;
  if (pharm_terms) *pharm_terms = s->pharm_terms;
  if (esp_terms) *esp_terms = s->esp_terms;
  return FSS_OK;
}

int fss_screen_admission(FssScreen* s, int64_t* pool_rows, int64_t* max_admitted, int64_t* memory_waits) {
  if (!s) return set_error(FSS_E_INVALID, "service is NULL");
  if (pool_rows)
    for (int c = 0; c < kQPools; ++c) pool_rows[c] = s->pool_rows[c];
  if (max_admitted) *max_admitted = s->admitted_max.load();
  if (memory_waits) *memory_waits = s->memory_waits.load();
  return FSS_OK;
}

int fss_screen_destroy(FssScreen* s) {
  if (!s) return FSS_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    DeviceGuard dg(s->lib->device);
    (void)hipStreamSynchronize(s->stream);
    free_arena(s);
  }
  {
    std::lock_guard<std::mutex> lk(s->lib->mu);
    s->lib->services -= 1;
  }
  delete s;
  return FSS_OK;
}

}  // extern "C"
