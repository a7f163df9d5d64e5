// Top-k merge of the screening service (internal to libfss_hip.so): the job table the service
// fills per launch and the launcher of the merge kernel.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace fss {

// One query of a launch: its current top-k (prev_n sorted entries) is merged with its k new pair
// scores (library indices base, base + 1, ...; entries off .. off + k - 1 of the launch's score
// and transform arrays) into out_* (every pointer is device memory).
struct MergeJob {
  const double* prev_s;
  const int64_t* prev_i;
  const float* prev_tf;
  double* out_s;
  int64_t* out_i;
  float* out_tf;
  int64_t base;
  int32_t prev_n, off, k, pad;
};

// Enqueues the merge of nj jobs (d_jobs in device memory) with K = top_k on `stream`; returns the
// launch status.
__attribute__((visibility("hidden"))) hipError_t topk_merge(int K, const MergeJob* d_jobs, int nj,
                                                             const double* new_s, const float* new_tf,
                                                             hipStream_t stream);

}  // namespace fss
