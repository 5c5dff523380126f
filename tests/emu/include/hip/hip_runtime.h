// CPU emulation of the HIP subset used by core_amd/csrc, for bounds checking
// under AddressSanitizer.  Each launch runs blockDim.x std::threads; each
// thread walks the blocks of the grid in order; __syncthreads is a block
// barrier, wave collectives use a per-wave barrier (convergent calls only).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <algorithm>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct alignas(16) int4 {
  int x, y, z, w;
};
extern thread_local dim3 threadIdx, blockIdx;
extern dim3 blockDim, gridDim;

namespace emu {
void launch(const char* where, dim3 grid, dim3 block, const std::function<void()>& fn);
void block_sync();
void lds_reg(void* p, size_t n);
uint64_t ballot(bool p);
int64_t shfl64(int64_t v, int src);
}  // namespace emu

inline void __syncthreads() { emu::block_sync(); }
inline uint64_t __ballot(int p) { return emu::ballot(p != 0); }
inline int __shfl(int v, int src) { return (int)emu::shfl64(v, src); }
inline double __shfl(double v, int src)
{
  int64_t b;
  memcpy(&b, &v, 8);
  b = emu::shfl64(b, src);
  double r;
  memcpy(&r, &b, 8);
  return r;
}
inline double __shfl_xor(double v, int mask)
{
  int64_t b;
  memcpy(&b, &v, 8);
  b = emu::shfl64(b, (int)((threadIdx.x & 63) ^ mask));
  double r;
  memcpy(&r, &b, 8);
  return r;
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }

inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v)
{
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline int atomicCAS(int* p, int cmp, int v)
{
  __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return cmp;
}
inline int atomicMax(int* p, int v)
{
  int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return old;
}

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNoDevice = 100 };
typedef void* hipStream_t;
struct emu_event { double t; };
typedef emu_event* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                     hipMemcpyDeviceToDevice, hipMemcpyDefault };
#define hipStreamNonBlocking 1

hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned f);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipMalloc(void** p, size_t n);
hipError_t hipFree(void* p);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st = 0);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st = 0);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
struct emu_graph;
typedef emu_graph* hipGraph_t;
typedef emu_graph* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = 0);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
