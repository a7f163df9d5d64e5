// Stand-in pair score of the sanitizer builds (test infrastructure): a deterministic function of
// a (query, library entry) pair's rows, used by hip_host.cpp in place of the GPU alignment and by
// screen_stress.cpp to predict every query's hits.
#pragma once
#include <math.h>
#include <stdint.h>

namespace standin {

inline double row_sum(const float* rows, int n, double wx, double wy, double wz) {
  double a = 0.0;
  for (int i = 0; i < n; ++i) a += wx * rows[4 * i] + wy * rows[4 * i + 1] + wz * rows[4 * i + 2];
  return a;
}

// Pair score in [0, 1] from the query's and the entry's shape rows and pharmacophore anchors and
// types (either side may be empty).
inline double score(const float* qpts, int nq, const float* epts, int ne, const float* qanc, const int32_t* qtyp,
                    int nqp, const float* eanc, const int32_t* etyp, int nep) {
  double a = row_sum(qpts, nq, 0.37, 0.11, 0.05) + row_sum(qanc, nqp, 0.23, 0.0, 0.13);
  double b = row_sum(epts, ne, 0.19, 0.0, -0.07) + row_sum(eanc, nep, 0.0, 0.31, 0.02);
  for (int i = 0; i < nqp; ++i) a += 0.01 * qtyp[i];
  for (int i = 0; i < nep; ++i) b += 0.017 * etyp[i];
  return 0.5 + 0.5 * sin(1.3 * a + 2.9 * b + 0.1 * ne);
}

// Steps the stand-in alignment reports for a pair of this score.
inline int32_t steps(double s) { return 1 + (int32_t)(s * 97.0) % 50; }

}  // namespace standin
