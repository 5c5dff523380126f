// fit_common.h -- pieces shared by the fast patch-fit kernels (k_fast.hip,
// k_tile.hip): the node-centred monomial basis, packed Cholesky solves and
// the route codes of a patch.
#ifndef SPRGPU_FIT_COMMON_H
#define SPRGPU_FIT_COMMON_H

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

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

// The Gram matrix of the 10-term basis from moments.  Its 55 entries
// sum_rows f_i f_j are sums of the 35 monomials x^a y^b z^c of degree <= 4
// of the centred coordinates, so each such moment is accumulated once: per
// row 3 additions and 31 fused multiply-adds of a coordinate or a
// degree-2 product, instead of 55.  kMomExp lists the moments in the order
// moments_row fills them; index 0 is the row count.
struct Exp3 {
  int a, b, c;
};
constexpr Exp3 kTermExp[10] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0},
                               {0, 1, 1}, {1, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2}};
constexpr int NMOM = 35;
constexpr Exp3 kMomExp[NMOM] = {
    {0, 0, 0},                                                                  // count
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},                                            // degree 1
    {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},           // degree 2
    {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1},
    {1, 0, 2}, {0, 1, 2}, {1, 1, 1},                                            // degree 3
    {4, 0, 0}, {0, 4, 0}, {0, 0, 4}, {3, 1, 0}, {3, 0, 1}, {1, 3, 0}, {0, 3, 1},
    {1, 0, 3}, {0, 1, 3}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1}, {1, 2, 1},
    {1, 1, 2}};                                                                 // degree 4

constexpr int mom_of(int a, int b, int c)
{
  for (int k = 0; k < NMOM; ++k)
    if (kMomExp[k].a == a && kMomExp[k].b == b && kMomExp[k].c == c) return k;
  return -1;
}

// The moment of packed Gram entry k = tri(i, j).
constexpr int gram_mom(int k)
{
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= k) ++i;
  const int j = k - i * (i + 1) / 2;
  return mom_of(kTermExp[i].a + kTermExp[j].a, kTermExp[i].b + kTermExp[j].b, kTermExp[i].c + kTermExp[j].c);
}
static_assert(gram_mom(0) == 0 && gram_mom(54) == mom_of(0, 0, 4) && gram_mom(tri(4, 4)) == mom_of(2, 2, 0),
              "Gram entries map to their moments");

// G[k] = M[gram_mom(k)] for the 55 packed entries, every index a
// compile-time constant (the moments stay in registers).
template <class T, int... K>
__device__ __forceinline__ void gram_from_moments_seq(const T* M, T* G, std::integer_sequence<int, K...>)
{
  ((G[K] = M[std::integral_constant<int, gram_mom(K)>::value]), ...);
}

template <class T>
__device__ __forceinline__ void gram_from_moments(const T* M, T* G)
{
  gram_from_moments_seq(M, G, std::make_integer_sequence<int, 55>{});
}

// Adds one row's moments 1..34 (centred coordinates d) to M.
template <class S>
__device__ __forceinline__ void moments_row(const double* d, S* M)
{
  const S x = (S)d[0], y = (S)d[1], z = (S)d[2];
  const S xy = x * y, yz = y * z, zx = z * x, xx = x * x, yy = y * y, zz = z * z;
  M[1] += x;
  M[2] += y;
  M[3] += z;
  M[4] = fma(x, y, M[4]);
  M[5] = fma(y, z, M[5]);
  M[6] = fma(z, x, M[6]);
  M[7] = fma(x, x, M[7]);
  M[8] = fma(y, y, M[8]);
  M[9] = fma(z, z, M[9]);
  M[10] = fma(x, xx, M[10]);
  M[11] = fma(y, yy, M[11]);
  M[12] = fma(z, zz, M[12]);
  M[13] = fma(y, xx, M[13]);
  M[14] = fma(z, xx, M[14]);
  M[15] = fma(x, yy, M[15]);
  M[16] = fma(z, yy, M[16]);
  M[17] = fma(x, zz, M[17]);
  M[18] = fma(y, zz, M[18]);
  M[19] = fma(x, yz, M[19]);
  M[20] = fma(xx, xx, M[20]);
  M[21] = fma(yy, yy, M[21]);
  M[22] = fma(zz, zz, M[22]);
  M[23] = fma(xx, xy, M[23]);
  M[24] = fma(xx, zx, M[24]);
  M[25] = fma(yy, xy, M[25]);
  M[26] = fma(yy, yz, M[26]);
  M[27] = fma(zz, zx, M[27]);
  M[28] = fma(zz, yz, M[28]);
  M[29] = fma(xx, yy, M[29]);
  M[30] = fma(xx, zz, M[30]);
  M[31] = fma(yy, zz, M[31]);
  M[32] = fma(xx, yz, M[32]);
  M[33] = fma(yy, zx, M[33]);
  M[34] = fma(zz, xy, M[34]);
}

}  // namespace sprg

#endif
