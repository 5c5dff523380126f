// k_mesh.hip -- upward adjacency of the device mesh mirror.
//
// getAdjacent(vertex|edge, 3) as CSR: a count pass, an exclusive scan, an
// atomic fill and a per-row sort.  Rows end up in ascending tet order, the
// order of the std::set<MeshEntity*> a patch is built in
// (spr/sprRecoverField.cc:110, handles of mds/apfMDS.cc:98-107), so the CSR
// is deterministic and bit-comparable with the reference adjacency.
#include "device.h"

namespace sprg {
namespace {

constexpr int SCAN_T = 256;   // threads per scan block
constexpr int SCAN_V = 8;     // values per thread
constexpr int SCAN_B = SCAN_T * SCAN_V;

__global__ void k_count(const int32_t* __restrict__ conn, int64_t n, int32_t* cnt,
                        int64_t nrows, int32_t* status)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t r = conn[i];
  if (r < 0 || r >= nrows) {
    atomicOr(status, 1);
    return;
  }
  atomicAdd(cnt + r + 1, 1);
}

// In-place inclusive scan of a[0..n) in blocks of SCAN_B; block totals to bsum.
__global__ __launch_bounds__(SCAN_T) void k_scan_block(int32_t* a, int64_t n, int32_t* bsum)
{
  __shared__ int32_t sh[SCAN_T];
  int64_t base = blockIdx.x * (int64_t)SCAN_B + threadIdx.x * (int64_t)SCAN_V;
  int32_t v[SCAN_V];
  int32_t s = 0;
  for (int k = 0; k < SCAN_V; ++k) {
    v[k] = (base + k < n) ? a[base + k] : 0;
    s += v[k];
    v[k] = s;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    int32_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int32_t pre = threadIdx.x ? sh[threadIdx.x - 1] : 0;
  for (int k = 0; k < SCAN_V; ++k)
    if (base + k < n) a[base + k] = v[k] + pre;
  if (threadIdx.x == SCAN_T - 1) bsum[blockIdx.x] = sh[SCAN_T - 1];
}

// Exclusive scan of the block totals by one block (nb is small).
__global__ __launch_bounds__(SCAN_T) void k_scan_bsum(int32_t* bsum, int64_t nb)
{
  __shared__ int32_t sh[SCAN_T];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nb; base += SCAN_T) {
    int64_t i = base + threadIdx.x;
    int32_t v = i < nb ? bsum[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
      int32_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == SCAN_T - 1) carry += sh[SCAN_T - 1];
    __syncthreads();
  }
}

__global__ void k_scan_add(int32_t* a, int64_t n, const int32_t* bsum)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[i] += bsum[i / SCAN_B];
}

__global__ void k_fill(const int32_t* __restrict__ conn, int64_t n, int per,
                       int64_t nrows, int32_t* cursor, int32_t* idx)
{
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t r = conn[i];
  if (r < 0 || r >= nrows) return;
  int32_t pos = atomicAdd(cursor + r, 1);
  idx[pos] = (int32_t)(i / per);
}

__global__ void k_sort_rows(const int32_t* __restrict__ off, int64_t nrows, int32_t* idx)
{
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int32_t b = off[r], e = off[r + 1];
  for (int32_t i = b + 1; i < e; ++i) {
    int32_t v = idx[i];
    int32_t j = i - 1;
    while (j >= b && idx[j] > v) {
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = v;
  }
}

inline unsigned grid(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

__global__ void k_copy_i32(int32_t* __restrict__ dst, const int32_t* __restrict__ src, int64_t n)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src ? src[i] : 0;
}

}  // namespace

// dst[0..n) = src[0..n), or 0 with src null.  A kernel rather than a
// memset/memcpy node, so that sequences captured into a HIP graph replay
// correctly.
int launch_copy_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s)
{
  if (n > 0) k_copy_i32<<<grid(n, 256), 256, 0, s>>>(dst, src, n);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "copy launch failed");
}

// In-place inclusive scan of a[0..n); work holds ceil(n/2048) int32.
int launch_scan_i32(int32_t* a, int64_t n, int32_t* work, hipStream_t s)
{
  if (n <= 0) return 0;
  int64_t nb = (n + SCAN_B - 1) / SCAN_B;
  k_scan_block<<<(unsigned)nb, SCAN_T, 0, s>>>(a, n, work);
  k_scan_bsum<<<1, SCAN_T, 0, s>>>(work, nb);
  k_scan_add<<<grid(n, 256), 256, 0, s>>>(a, n, work);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "scan launch failed");
}

// tmp must hold max(nrows, ceil(nrows/SCAN_B)) + 2 int32.
int launch_csr(const MeshView& m, int dim, int32_t* off, int32_t* idx,
               int32_t* tmp, hipStream_t s)
{
  const int32_t* conn = dim == 0 ? m.tv : m.te;
  int per = dim == 0 ? 4 : 6;
  int64_t nrows = dim == 0 ? m.nv : m.ne;
  int64_t n = per * m.nt;
  int32_t* status = tmp;       // 1 int
  int32_t* work = tmp + 2;     // cursor / block sums
  if (hipMemsetAsync(off, 0, sizeof(int32_t) * (nrows + 1), s) != hipSuccess ||
      hipMemsetAsync(status, 0, sizeof(int32_t), s) != hipSuccess)
    return sprg_fail(SPRG_EHIP, "csr: memset failed");
  if (n > 0) k_count<<<grid(n, 256), 256, 0, s>>>(conn, n, off, nrows, status);
  // inclusive scan of off[1..nrows] gives the exclusive offsets
  int64_t nb = (nrows + SCAN_B - 1) / SCAN_B;
  if (nrows > 0) {
    k_scan_block<<<(unsigned)nb, SCAN_T, 0, s>>>(off + 1, nrows, work);
    k_scan_bsum<<<1, SCAN_T, 0, s>>>(work, nb);
    k_scan_add<<<grid(nrows, 256), 256, 0, s>>>(off + 1, nrows, work);
  }
  if (hipMemcpyAsync(work, off, sizeof(int32_t) * nrows, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return sprg_fail(SPRG_EHIP, "csr: copy failed");
  if (n > 0) k_fill<<<grid(n, 256), 256, 0, s>>>(conn, n, per, nrows, work, idx);
  if (nrows > 0) k_sort_rows<<<grid(nrows, 256), 256, 0, s>>>(off, nrows, idx);
  int32_t st = 0;
  if (hipMemcpyAsync(&st, status, sizeof st, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return sprg_fail(SPRG_EHIP, "csr: %s", hipGetErrorString(hipGetLastError()));
  if (st) return sprg_fail(SPRG_EINVAL, "mesh: connectivity index out of range");
  return 0;
}

}  // namespace sprg
