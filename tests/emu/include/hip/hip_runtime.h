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
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) ((const void*)&(x))

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
void launch(const char* where, dim3 grid, dim3 block, size_t shmem, const std::function<void()>& fn);
// the launch's dynamic LDS (extern __shared__), shmem bytes, poisoned like
// static LDS before each block
void* dyn_lds();
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
inline float __shfl_xor(float v, int mask)
{
  int32_t b;
  memcpy(&b, &v, 4);
  b = (int32_t)emu::shfl64(b, (int)((threadIdx.x & 63) ^ mask));
  float r;
  memcpy(&r, &b, 4);
  return r;
}
// DPP quad permutation (dpp_ctrl < 0x100) with every row and bank enabled;
// other controls are not emulated
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound)
{
  (void)old;
  (void)bound;
  if (ctrl >= 0x100 || row_mask != 0xF || bank_mask != 0xF) {
    fprintf(stderr, "EMU: DPP control 0x%x not emulated\n", ctrl);
    abort();
  }
  const int l = threadIdx.x & 63;
  const int from = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
  return (int)emu::shfl64(src, from);
}
// the value of lane `lane` of the wave (every lane calls it)
inline int __builtin_amdgcn_readlane(int src, int lane) { return (int)emu::shfl64(src, lane); }
inline int __builtin_amdgcn_mov_dpp(int src, int ctrl, int row_mask, int bank_mask, bool bound)
{
  return __builtin_amdgcn_update_dpp(0, src, ctrl, row_mask, bank_mask, bound);
}
// the hardware reciprocal square root estimates, exact here (the callers
// refine them with Newton steps)
inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }
inline float __builtin_amdgcn_rsqf(float x) { return 1.0f / std::sqrt(x); }
// ds_swizzle in bitmask mode (offset bit 15 clear): within groups of 32
// lanes, lane l reads ((l & and) | or) ^ xor
inline int __builtin_amdgcn_ds_swizzle(int src, int pattern)
{
  if (pattern & 0x8000) {
    fprintf(stderr, "EMU: swizzle mode 0x%x not emulated\n", pattern);
    abort();
  }
  const int and_m = pattern & 0x1F, or_m = (pattern >> 5) & 0x1F, xor_m = (pattern >> 10) & 0x1F;
  const int l = threadIdx.x & 63;
  const int from = (l & 32) | (((l & 31 & and_m) | or_m) ^ xor_m);
  return (int)emu::shfl64(src, from);
}
inline int __double2loint(double x) { uint64_t u; memcpy(&u, &x, 8); return (int)(uint32_t)u; }
inline int __double2hiint(double x) { uint64_t u; memcpy(&u, &x, 8); return (int)(uint32_t)(u >> 32); }
inline double __hiloint2double(int hi, int lo)
{
  const uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
  double x;
  memcpy(&x, &u, 8);
  return x;
}
inline int __float_as_int(float x) { int i; memcpy(&i, &x, 4); return i; }
inline float __int_as_float(int i) { float x; memcpy(&x, &i, 4); return x; }
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
// device arrays are host arrays here: the symbol's address is the array's
inline hipError_t hipGetSymbolAddress(void** p, const void* sym)
{
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
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
