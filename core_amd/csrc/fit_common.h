// fit_common.h -- pieces shared by the fast patch-fit kernels (k_fast.hip,
// k_tile.hip): the node-centred monomial basis, packed Cholesky solves and
// the route codes of a patch.
#ifndef SPRGPU_FIT_COMMON_H
#define SPRGPU_FIT_COMMON_H

#include <hip/hip_runtime.h>

#include "device.h"

namespace sprg {

// What a fit pass decides for a patch.  ACT_EXACT: the exact path decides
// the rank (and expands); ACT_FALLBACK: an FP64 pass refits it; ACT_REFIT:
// the exact kernel refits the certified set with the reference's
// arithmetic; ACT_BIG: the same for a set above 192 rows (k_exact_big);
// ACT_ILLCOND: no form here can fit it.
enum { ACT_SOLVE = 0, ACT_EXACT = 1, ACT_FALLBACK = 2, ACT_ILLCOND = 3, ACT_NONE = 4, ACT_REFIT = 5, ACT_BIG = 6 };

// Appends v to the big list (count: status[3]); every lane of the wave
// calls it.
__device__ __forceinline__ void big_append(bool mine, const Lists& L, int32_t v)
{
  wave_append(1, mine ? 1 : 0, L.status + 3, L.big, v);
}

template <class S, int NT>
__device__ __forceinline__ void basis(const double* d, S* f)
{
  const S a = (S)d[0], b = (S)d[1], c = (S)d[2];
  f[0] = 1;
  f[1] = a;
  f[2] = b;
  f[3] = c;
  if (NT == 10) {
    f[4] = a * b;
    f[5] = b * c;
    f[6] = c * a;
    f[7] = a * a;
    f[8] = b * b;
    f[9] = c * c;
  }
}

// packed lower triangle index
__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }

template <class T, int NT>
__device__ __forceinline__ void chol_solve(const T* L, const T* rhs, T* z)
{
  T y[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    T s = rhs[k];
#pragma unroll
    for (int j = 0; j < k; ++j) s = fma(-L[tri(k, j)], y[j], s);
    y[k] = s / L[tri(k, k)];
  }
#pragma unroll
  for (int kk = 0; kk < NT; ++kk) {
    const int k = NT - 1 - kk;
    T s = y[k];
#pragma unroll
    for (int j = k + 1; j < NT; ++j) s = fma(-L[tri(j, k)], z[j], s);
    z[k] = s / L[tri(k, k)];
  }
}

}  // namespace sprg

#endif
