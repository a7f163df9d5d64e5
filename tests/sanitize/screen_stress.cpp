// Drives the screening service's host code (fss_screen.cpp linked with the stand-ins of
// hip_host.cpp) from several threads, for the sanitizer builds of tests/sanitize/build.py (test
// infrastructure):
//   * four threads submit queries of random sizes while three threads poll them (waiting without
//     limit, with a timeout, or not at all, and retrying PENDING tickets), a fifth thread appends
//     conformers to the library mid-run, and the row pools are sized so queries wait for memory;
//   * every query's hits must be the top-k that the stand-in scores predict for the library size
//     at its admission (exactly known for queries that ask for every entry's score);
//   * a query on an empty library completes with no hits, an oversized one is refused, and a
//     service closed with queries still queued shuts down cleanly.
// Modes: vol (shape rows) and vol_color (shape rows + pharmacophores, several pairs per
// workgroup, so each launch is also ordered by entry size).  Exit status 0 on success.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../include/fss.h"
#include "standin.h"

namespace {

int g_failures = 0;
std::mutex g_out;

void fail(const char* what, long a = 0, long b = 0) {
  std::lock_guard<std::mutex> lk(g_out);
  fprintf(stderr, "FAIL: %s (%ld, %ld)\n", what, a, b);
  ++g_failures;
}

struct Rows {   // conformers in fss_lib_append's layout
  std::vector<float> pts, anc, vec;
  std::vector<int32_t> cnt, typ, pcnt;
};

Rows make_rows(std::mt19937& g, int n, int max_pts, int max_ph, int min_pts = 1) {
  std::uniform_real_distribution<float> u(-3.f, 3.f);
  std::uniform_int_distribution<int> np(min_pts, max_pts), nph(1, max_ph), ty(0, 7);
  Rows r;
  for (int i = 0; i < n; ++i) {
    const int k = np(g), m = nph(g);
    r.cnt.push_back(k);
    r.pcnt.push_back(m);
    for (int j = 0; j < 4 * k; ++j) r.pts.push_back((j & 3) == 3 ? 0.f : u(g));
    for (int j = 0; j < 4 * m; ++j) {
      r.anc.push_back((j & 3) == 3 ? 0.f : u(g));
      r.vec.push_back((j & 3) == 3 ? 0.f : u(g));
    }
    for (int j = 0; j < m; ++j) r.typ.push_back(ty(g));
  }
  return r;
}

// Host view of the library as appended (entry i: rows from the offsets).
struct HostLib {
  std::vector<float> pts, anc;
  std::vector<int32_t> typ;
  std::vector<int64_t> off, n, poff, pn;
  void add(const Rows& r) {
    int64_t o = (int64_t)pts.size() / 4, po = (int64_t)typ.size();
    for (size_t i = 0; i < r.cnt.size(); ++i) {
      off.push_back(o); n.push_back(r.cnt[i]); o += r.cnt[i];
      poff.push_back(po); pn.push_back(r.pcnt[i]); po += r.pcnt[i];
    }
    pts.insert(pts.end(), r.pts.begin(), r.pts.end());
    anc.insert(anc.end(), r.anc.begin(), r.anc.end());
    typ.insert(typ.end(), r.typ.begin(), r.typ.end());
  }
};

struct Query {
  std::vector<float> pts, anc, vec;
  std::vector<int32_t> typ;
  bool all = false;
  bool small_polled = false;   // an all-scores query polled once with too small a buffer
  int64_t ticket = 0;
};

double expect_score(const Query& q, const HostLib& L, int64_t e, bool pharm) {
  return standin::score(q.pts.data(), (int)q.pts.size() / 4, L.pts.data() + 4 * L.off[e], (int)L.n[e],
                        pharm ? q.anc.data() : nullptr, pharm ? q.typ.data() : nullptr,
                        pharm ? (int)q.typ.size() : 0, pharm ? L.anc.data() + 4 * L.poff[e] : nullptr,
                        pharm ? L.typ.data() + L.poff[e] : nullptr, pharm ? (int)L.pn[e] : 0);
}

// The hits of q over entries [0, size): (score desc, index asc), top K.
std::vector<std::pair<double, int64_t>> expect_hits(const Query& q, const HostLib& L, int64_t size, int K, bool pharm) {
  std::vector<std::pair<double, int64_t>> h;
  for (int64_t e = 0; e < size; ++e) h.push_back({expect_score(q, L, e, pharm), e});
  std::sort(h.begin(), h.end(), [](const auto& a, const auto& b) {
    return a.first > b.first || (a.first == b.first && a.second < b.second);
  });
  if ((int)h.size() > K) h.resize(K);
  return h;
}

bool same_hits(const FssHits& got, const std::vector<std::pair<double, int64_t>>& want) {
  if (got.n != (int)want.size()) return false;
  for (int i = 0; i < got.n; ++i) {
    if (got.score[i] != want[i].first || got.index[i] != want[i].second) return false;
    for (int k = 0; k < 16; ++k)
      if (got.transform[16 * i + k] != (float)want[i].first + (float)k) return false;
  }
  return true;
}

void run_mode(int mode, unsigned seed) {
  const bool pharm = mode == FSS_MODE_SHAPE_PHARM;
  const int n0 = 1500, n1 = 500, K = 25, max_pts = 12, max_ph = 6;
  std::mt19937 g(seed);
  const Rows r0 = make_rows(g, n0, max_pts, max_ph), r1 = make_rows(g, n1, max_pts, max_ph);
  HostLib host;
  host.add(r0);
  host.add(r1);

  FssLibSpec spec{};
  spec.channels = FSS_LIB_SHAPE | FSS_LIB_PHARM;
  spec.max_conformers = n0 + n1;
  spec.max_points = (int64_t)(r0.pts.size() + r1.pts.size()) / 4;
  spec.max_pharm = (int64_t)(r0.typ.size() + r1.typ.size());
  FssLibrary* lib = nullptr;
  if (fss_lib_create(&spec, &lib)) { fail("fss_lib_create"); return; }
  auto append = [&](const Rows& r) {
    FssLibRows in{};
    in.pts = r.pts.data(); in.counts = r.cnt.data();
    in.anc = r.anc.data(); in.vec = r.vec.data(); in.typ = r.typ.data(); in.pcounts = r.pcnt.data();
    int64_t first = -1;
    if (fss_lib_append(lib, (int32_t)r.cnt.size(), &in, &first)) fail("fss_lib_append");
  };
  append(r0);

  FssScoreParams sp{};
  sp.mode = mode;
  sp.alpha = 0.81;
  sp.lam = 1.0;
  const float quats[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  FssScreenParams prm{};
  prm.n_restarts = 4; prm.max_steps = 10; prm.lr = 0.1f; prm.beta1 = 0.9f; prm.beta2 = 0.999f; prm.eps = 1e-8f;
  prm.early_stop = 1; prm.es_tol = 1e-5f; prm.es_patience = 10;
  prm.top_k = K; prm.launch_pairs = 700; prm.max_active = 3;
  prm.max_query_points = max_pts; prm.max_query_pharm = max_ph;
  prm.quat_table = quats;
  // room for about two maximum queries with three slots: large queries wait for memory
  prm.query_budget_bytes = pharm ? 2 * (16 * max_pts + 36 * max_ph) + 100 : 2 * 16 * max_pts + 40;
  prm.pairs_per_wg = pharm ? 3 : 0;
  FssScreen* svc = nullptr;
  if (fss_screen_create(lib, &sp, &prm, &svc)) { fail("fss_screen_create"); fss_lib_destroy(lib); return; }

  const int n_submit = 4, per = 12;
  std::vector<Query> queries(n_submit * per);
  for (size_t i = 0; i < queries.size(); ++i) {
    // every other query of the maximum size: the pools hold two, so queries wait for room
    const Rows one = make_rows(g, 1, max_pts, max_ph, i % 2 ? max_pts : 1);
    queries[i].pts = one.pts;
    queries[i].anc = one.anc;
    queries[i].vec = one.vec;
    queries[i].typ = one.typ;
    queries[i].all = i % 5 == 0;
  }
  std::mutex qmu;
  std::deque<int> ready;                 // submitted, not yet collected
  std::atomic<int> submitted{0}, collected{0};
  std::vector<std::thread> th;
  for (int w = 0; w < n_submit; ++w) {
    th.emplace_back([&, w] {
      for (int k = 0; k < per; ++k) {
        Query& q = queries[w * per + k];
        FssQuery fq{};
        fq.n_points = (int32_t)q.pts.size() / 4;
        fq.pts = q.pts.data();
        if (pharm) {
          fq.n_pharm = (int32_t)q.typ.size();
          fq.anc = q.anc.data(); fq.vec = q.vec.data(); fq.typ = q.typ.data();
        }
        fq.all_scores = q.all;
        int64_t ticket = 0;
        if (fss_screen_submit(svc, &fq, &ticket)) { fail("submit", w, k); continue; }
        q.ticket = ticket;
        submitted.fetch_add(1);
        std::lock_guard<std::mutex> lk(qmu);
        ready.push_back(w * per + k);
      }
    });
  }
  th.emplace_back([&] {   // the library grows while queries run
    while (submitted.load() < 8) std::this_thread::yield();
    append(r1);
  });
  const int total = n_submit * per;
  for (int w = 0; w < 3; ++w) {
    th.emplace_back([&, w] {
      std::vector<double> s(K), all(n0 + n1);
      std::vector<int64_t> idx(K);
      std::vector<float> tf(16 * K);
      int turn = 0;
      while (collected.load() < total) {
        int qi = -1;
        {
          std::lock_guard<std::mutex> lk(qmu);
          if (!ready.empty()) { qi = ready.front(); ready.pop_front(); }
        }
        if (qi < 0) { std::this_thread::yield(); continue; }
        Query& q = queries[qi];
        FssHits h{};
        h.capacity = K; h.score = s.data(); h.index = idx.data(); h.transform = tf.data();
        h.all_scores = all.data();
        // an all-scores query polled with too small a buffer keeps its result held for a retry
        const bool small = q.all && !q.small_polled;
        h.n_all = small ? 1 : (int64_t)all.size();
        const int pick = turn++ + w;
        const int wait = pick % 3 == 0 ? -1 : (pick % 3 == 1 ? 0 : 2);
        const int rc = fss_screen_poll(svc, q.ticket, wait, &h);
        if (small && rc == FSS_E_INVALID) q.small_polled = true;
        if (rc == FSS_PENDING || (small && rc == FSS_E_INVALID)) {
          std::lock_guard<std::mutex> lk(qmu);
          ready.push_back(qi);
          continue;
        }
        collected.fetch_add(1);
        if (rc) { fail("poll", (long)q.ticket, rc); continue; }
        bool ok;
        if (q.all) {   // the library size at admission is known: compare every entry
          const int64_t L = h.n_all;
          ok = (L == n0 || L == n0 + n1) && same_hits(h, expect_hits(q, host, L, K, pharm));
          for (int64_t e = 0; ok && e < L; ++e) ok = all[e] == expect_score(q, host, e, pharm);
        } else {
          ok = same_hits(h, expect_hits(q, host, n0, K, pharm)) || same_hits(h, expect_hits(q, host, n0 + n1, K, pharm));
        }
        if (!ok) fail("hits differ from the stand-in prediction", (long)q.ticket, mode);
        if (h.latency_ms < h.queue_ms || h.queue_ms < 0) fail("latency < queue time", (long)q.ticket);
      }
    });
  }
  for (auto& t : th) t.join();
  int64_t launches = 0, pairs = 0, steps = 0, pool_rows[FSS_QPOOLS], max_adm = 0, waits = 0;
  fss_screen_stats(svc, &launches, &pairs, &steps);
  fss_screen_admission(svc, pool_rows, &max_adm, &waits);
  if (launches < 1 || pairs < (int64_t)total * n0 || steps < pairs) fail("stats", (long)launches, (long)pairs);
  if (waits < 1) fail("no query waited for pool room", (long)waits);
  printf("mode %d: %d queries, %lld launches, %lld pairs, max admitted %lld, memory waits %lld\n", mode, total,
         (long long)launches, (long long)pairs, (long long)max_adm, (long long)waits);

  // oversized query: refused at submit
  {
    std::vector<float> big(4 * (max_pts + 1), 0.f);
    FssQuery fq{};
    fq.n_points = max_pts + 1; fq.pts = big.data();
    if (pharm) { fq.n_pharm = (int32_t)queries[0].typ.size(); fq.anc = queries[0].anc.data(); fq.vec = queries[0].vec.data(); fq.typ = queries[0].typ.data(); }
    int64_t ticket = 0;
    if (fss_screen_submit(svc, &fq, &ticket) != FSS_E_TOO_LARGE) fail("oversized query accepted");
  }
  // closing with queries queued and running
  for (int k = 0; k < 9; ++k) {
    const Query& q = queries[k];
    FssQuery fq{};
    fq.n_points = (int32_t)q.pts.size() / 4; fq.pts = q.pts.data();
    if (pharm) { fq.n_pharm = (int32_t)q.typ.size(); fq.anc = q.anc.data(); fq.vec = q.vec.data(); fq.typ = q.typ.data(); }
    int64_t ticket = 0;
    if (fss_screen_submit(svc, &fq, &ticket)) fail("submit before close", k);
  }
  if (fss_screen_destroy(svc)) fail("fss_screen_destroy");
  if (fss_lib_destroy(lib)) fail("fss_lib_destroy");
}

void run_empty_library() {
  FssLibSpec spec{};
  spec.channels = FSS_LIB_SHAPE;
  spec.max_conformers = 4;
  spec.max_points = 16;
  FssLibrary* lib = nullptr;
  if (fss_lib_create(&spec, &lib)) { fail("empty fss_lib_create"); return; }
  FssScoreParams sp{};
  sp.mode = FSS_MODE_SHAPE; sp.alpha = 0.81; sp.lam = 1.0;
  const float quats[4] = {1, 0, 0, 0};
  FssScreenParams prm{};
  prm.n_restarts = 1; prm.max_steps = 5; prm.lr = 0.1f; prm.beta1 = 0.9f; prm.beta2 = 0.999f; prm.eps = 1e-8f;
  prm.top_k = 5; prm.launch_pairs = 8; prm.max_active = 2; prm.max_query_points = 4; prm.quat_table = quats;
  FssScreen* svc = nullptr;
  if (fss_screen_create(lib, &sp, &prm, &svc)) { fail("empty fss_screen_create"); fss_lib_destroy(lib); return; }
  const float pts[8] = {0, 0, 0, 0, 1, 0, 0, 0};
  FssQuery fq{};
  fq.n_points = 2; fq.pts = pts;
  int64_t ticket = 0;
  double s[5];
  int64_t idx[5];
  float tf[80];
  FssHits h{};
  h.capacity = 5; h.score = s; h.index = idx; h.transform = tf;
  if (fss_screen_submit(svc, &fq, &ticket) || fss_screen_poll(svc, ticket, -1, &h) || h.n != 0)
    fail("empty library query", h.n);
  if (fss_screen_poll(svc, ticket, 0, &h) != FSS_E_INVALID) fail("a collected ticket polled again");
  fss_screen_destroy(svc);
  fss_lib_destroy(lib);
}

}  // namespace

int main() {
  run_mode(FSS_MODE_SHAPE, 11);
  run_mode(FSS_MODE_SHAPE_PHARM, 12);
  run_empty_library();
  if (g_failures) {
    fprintf(stderr, "%d failures\n", g_failures);
    return 1;
  }
  printf("screen stress: OK\n");
  return 0;
}
