// Top-k merge kernel of the screening service (see fss_screen.cpp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fss.h"
#include "fss_topk.h"

namespace fss {

namespace {

// ------------------------------------------------------------------------------------------------
// Top-k merge: one workgroup per query of a launch.  Candidates are the query's current top-k
// (prev_n entries, sorted) followed by its k new pair scores (library indices base, base+1, ...).
// The K-th largest key is found by an 8-bit radix select over order-preserving 64-bit keys of the
// FP64 scores; everything above it and the first ties in candidate order (which is library-index
// order among equal scores) are kept, then sorted by (score desc, index asc) with a bitonic sort.
// ------------------------------------------------------------------------------------------------
constexpr int kMergeThreads = 1024;

__device__ __forceinline__ uint64_t order_key(double v) {
  if (v != v) return 0ull;  // NaN sorts last
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__global__ __launch_bounds__(kMergeThreads) void k_topk_merge(int K, const MergeJob* __restrict__ jobs,
                                                              const double* __restrict__ new_s,
                                                              const float* __restrict__ new_tf) {
  __shared__ uint32_t hist[256];
  __shared__ uint64_t skey[FSS_SCREEN_MAX_TOPK];
  __shared__ int64_t sidx[FSS_SCREEN_MAX_TOPK];
  __shared__ int32_t scand[FSS_SCREEN_MAX_TOPK];
  __shared__ uint32_t wave_cnt[kMergeThreads / 64];
  __shared__ uint64_t s_prefix;
  __shared__ int32_t s_count, s_ties;
  const MergeJob jb = jobs[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = jb.prev_n + jb.k;
  const int need = min(K, n);
  auto key_of = [&](int c) -> uint64_t {
    return order_key(c < jb.prev_n ? jb.prev_s[c] : new_s[jb.off + (c - jb.prev_n)]);
  };
  auto idx_of = [&](int c) -> int64_t { return c < jb.prev_n ? jb.prev_i[c] : jb.base + (c - jb.prev_n); };
  if (tid == 0) { s_count = 0; s_ties = 0; }
  __syncthreads();
  if (n <= K) {
    for (int c = tid; c < n; c += nt) { skey[c] = key_of(c); sidx[c] = idx_of(c); scand[c] = c; }
  } else {
    // radix select of the K-th largest key
    uint64_t prefix = 0, mask = 0;
    int remaining = K;
    for (int shift = 56; shift >= 0; shift -= 8) {
      for (int i = tid; i < 256; i += nt) hist[i] = 0u;
      __syncthreads();
      for (int c = tid; c < n; c += nt) {
        const uint64_t key = key_of(c);
        if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
      }
      __syncthreads();
      if (tid == 0) {
        int cum = 0, d = 255;
        for (; d > 0; --d) {
          if (cum + (int)hist[d] >= remaining) break;
          cum += (int)hist[d];
        }
        s_count = remaining - cum;          // ties still needed inside digit d
        s_prefix = prefix | ((uint64_t)d << shift);
      }
      __syncthreads();
      remaining = s_count;
      prefix = s_prefix;
      mask |= 0xFFull << shift;
      __syncthreads();
    }
    const uint64_t T = prefix;              // the K-th largest key; `remaining` copies of it are kept
    if (tid == 0) s_count = 0;
    __syncthreads();
    for (int c = tid; c < n; c += nt) {
      const uint64_t key = key_of(c);
      if (key > T) {
        const int pos = atomicAdd(&s_count, 1);
        skey[pos] = key; sidx[pos] = idx_of(c); scand[pos] = c;
      }
    }
    __syncthreads();
    const int above = s_count;              // == K - remaining
    for (int c0 = 0; c0 < n; c0 += nt) {
      const int c = c0 + tid;
      const bool tie = c < n && key_of(c) == T;
      const uint64_t bal = __ballot(tie);
      const int lane = tid & 63, w = tid >> 6;
      const int in_wave = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) wave_cnt[w] = (uint32_t)__popcll(bal);
      __syncthreads();
      int before = s_ties;
      for (int i = 0; i < w; ++i) before += (int)wave_cnt[i];
      const int pos = before + in_wave;
      if (tie && pos < remaining) { skey[above + pos] = T; sidx[above + pos] = idx_of(c); scand[above + pos] = c; }
      __syncthreads();
      if (tid == 0) {
        int tot = 0;
        for (int i = 0; i < nt / 64; ++i) tot += (int)wave_cnt[i];
        s_ties += tot;
      }
      __syncthreads();
      if (s_ties >= remaining) break;
    }
  }
  __syncthreads();
  // bitonic sort of the `need` selected entries by (key desc, index asc); padding sorts last
  int L = 1;
  while (L < need) L <<= 1;
  for (int i = need + tid; i < L; i += nt) { skey[i] = 0ull; sidx[i] = INT64_MAX; scand[i] = -1; }
  __syncthreads();
  for (int size = 2; size <= L; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < L; i += nt) {
        const int j = i ^ stride;
        if (j > i) {
          // "before" = larger key, then smaller index
          const bool i_first = skey[i] > skey[j] || (skey[i] == skey[j] && sidx[i] < sidx[j]);
          const bool up = (i & size) == 0;
          if (up != i_first) {
            const uint64_t tk = skey[i]; skey[i] = skey[j]; skey[j] = tk;
            const int64_t ti = sidx[i]; sidx[i] = sidx[j]; sidx[j] = ti;
            const int32_t tc = scand[i]; scand[i] = scand[j]; scand[j] = tc;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int r = tid; r < need; r += nt) {
    const int c = scand[r];
    const bool prev = c < jb.prev_n;
    jb.out_s[r] = prev ? jb.prev_s[c] : new_s[jb.off + (c - jb.prev_n)];
    jb.out_i[r] = sidx[r];
  }
  for (int e = tid; e < need * 16; e += nt) {
    const int r = e >> 4, m = e & 15;
    const int c = scand[r];
    jb.out_tf[e] = c < jb.prev_n ? jb.prev_tf[16 * c + m] : new_tf[16 * (jb.off + (c - jb.prev_n)) + m];
  }
}

}  // namespace

hipError_t topk_merge(int K, const MergeJob* d_jobs, int nj, const double* new_s, const float* new_tf,
                      hipStream_t stream) {
  hipLaunchKernelGGL(k_topk_merge, dim3(nj), dim3(kMergeThreads), 0, stream, K, d_jobs, new_s, new_tf);
  return hipGetLastError();
}

}  // namespace fss
