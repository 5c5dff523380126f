// Launch machinery of the CPU emulation (see include/hip/hip_runtime.h).
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <climits>
#include <cstdint>
#include <mutex>
#include <thread>
#include <vector>

#include <linux/futex.h>
#include <sys/syscall.h>
#include <time.h>
#include <unistd.h>

thread_local dim3 threadIdx, blockIdx;

dim3 blockDim, gridDim;

namespace emu {
namespace {

// A barrier of n threads: an arrival counter and a generation word.  The
// last thread to arrive resets the counter, bumps the generation and wakes
// the sleepers with one futex call; the others spin briefly, then sleep on
// the generation word (no mutex for every woken thread to take in turn).
struct Barrier {
  std::atomic<int> count{0};
  std::atomic<uint32_t> gen{0};
  int n = 0;
  void reset(int c)
  {
    n = c;
    count.store(0);
  }
  void wait(const char* what)
  {
    const uint32_t g = gen.load(std::memory_order_acquire);
    if (count.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
      count.store(0, std::memory_order_relaxed);
      gen.fetch_add(1, std::memory_order_acq_rel);
      syscall(SYS_futex, reinterpret_cast<uint32_t*>(&gen), FUTEX_WAKE_PRIVATE, INT32_MAX, nullptr, nullptr, 0);
      return;
    }
    for (int i = 0; i < 64; ++i) {
      if (gen.load(std::memory_order_acquire) != g) return;
      std::this_thread::yield();
    }
    const auto t0 = std::chrono::steady_clock::now();
    while (gen.load(std::memory_order_acquire) == g) {
      struct timespec ts = {1, 0};
      syscall(SYS_futex, reinterpret_cast<uint32_t*>(&gen), FUTEX_WAIT_PRIVATE, g, &ts, nullptr, 0);
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) {
        fprintf(stderr, "EMU DEADLOCK at %s (block %u thread %u, %d of %d waiting)\n", what,
                blockIdx.x, threadIdx.x, count.load(), n);
        abort();
      }
    }
  }
};

const char* g_where = "";
Barrier g_block;
std::vector<Barrier*> g_wave;
int64_t g_slots[1024];

}  // namespace

static std::mutex g_lds_mu;
static std::vector<std::pair<void*, size_t>> g_lds;
static bool g_poison = getenv("EMU_POISON") != nullptr;

void lds_reg(void* p, size_t n)
{
  std::lock_guard<std::mutex> lk(g_lds_mu);
  for (auto& r : g_lds)
    if (r.first == p) return;
  g_lds.push_back({p, n});
  if (g_poison) memset(p, 0xFF, n);
}

static void lds_poison_all()
{
  std::lock_guard<std::mutex> lk(g_lds_mu);
  for (auto& r : g_lds) memset(r.first, 0xFF, r.second);
}

void block_sync() { g_block.wait(g_where); }

uint64_t ballot(bool p)
{
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  g_slots[threadIdx.x] = p;
  g_wave[w]->wait("ballot");
  uint64_t m = 0;
  const int nl = std::min<int>(64, (int)blockDim.x - 64 * w);
  for (int i = 0; i < nl; ++i)
    if (g_slots[64 * w + i]) m |= 1ull << i;
  g_wave[w]->wait("ballot2");
  (void)lane;
  return m;
}

int64_t shfl64(int64_t v, int src)
{
  const int w = threadIdx.x / 64;
  g_slots[threadIdx.x] = v;
  g_wave[w]->wait("shfl");
  int64_t r = g_slots[64 * w + (src & 63)];
  g_wave[w]->wait("shfl2");
  return r;
}

struct Rec {
  const char* where;
  dim3 grid, block;
  size_t shmem;
  std::function<void()> fn;
};

// dynamic LDS of the running launch: a heap block of exactly the requested
// size, so that AddressSanitizer reports any access past it
static void* g_dyn = nullptr;
static size_t g_dyn_bytes = 0;
void* dyn_lds() { return g_dyn; }
static bool g_capturing = false;
static std::vector<Rec> g_rec;

void run_launch(const char* where, dim3 grid, dim3 block, size_t shmem, const std::function<void()>& fn);

void launch(const char* where, dim3 grid, dim3 block, size_t shmem, const std::function<void()>& fn)
{
  if (g_capturing) {
    g_rec.push_back({where, grid, block, shmem, fn});
    return;
  }
  run_launch(where, grid, block, shmem, fn);
}

void run_launch(const char* where, dim3 grid, dim3 block, size_t shmem, const std::function<void()>& fn)
{
  if (shmem > 64 * 1024) {
    fprintf(stderr, "EMU launch %s: %zu bytes of dynamic LDS exceed 64 KB\n", where, shmem);
    abort();
  }
  free(g_dyn);
  g_dyn = shmem ? malloc(shmem) : nullptr;
  g_dyn_bytes = shmem;
  if (block.x == 0 || block.x > 1024 || block.y != 1 || block.z != 1 || grid.y != 1 || grid.z != 1) {
    fprintf(stderr, "EMU bad launch %s: grid %u block %u\n", where, grid.x, block.x);
    abort();
  }
  if (grid.x == 0) return;
  // persistent grid-stride kernels: a few blocks cover the same work
  if (!strncmp(where, "k_exact_fit", 11) || !strncmp(where, "k_expand", 8) || !strncmp(where, "k_fast", 6)) {
    static int cap = getenv("EMU_GRID") ? atoi(getenv("EMU_GRID")) : 6;
    if ((int)grid.x > cap) grid.x = cap;
  }
  g_where = where;
  blockDim = block;
  gridDim = grid;
  g_block.reset(block.x);
  int nw = (block.x + 63) / 64;
  while ((int)g_wave.size() < nw) g_wave.push_back(new Barrier);
  for (int w = 0; w < nw; ++w) g_wave[w]->reset(std::min<int>(64, block.x - 64 * w));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < block.x; ++t)
    th.emplace_back([&, t] {
      threadIdx = dim3(t);
      for (unsigned b = 0; b < grid.x; ++b) {
        blockIdx = dim3(b);
        if (g_poison) {
          // fresh LDS holds garbage: fill every __shared__ variable of the
          // kernels run so far with 0xFF (int -1, double NaN) before a block
          if (t == 0) {
            lds_poison_all();
            if (g_dyn) memset(g_dyn, 0xFF, g_dyn_bytes);
          }
          g_block.wait("poison");
        }
        fn();
        g_block.wait("end of block");
      }
    });
  for (auto& t : th) t.join();
}

bool capturing() { return g_capturing; }
void begin_capture() { g_capturing = true; g_rec.clear(); }
std::vector<Rec> end_capture() { g_capturing = false; std::vector<Rec> r; r.swap(g_rec); return r; }
void replay(const std::vector<Rec>& r) { for (auto& x : r) run_launch(x.where, x.grid, x.block, x.shmem, x.fn); }

}  // namespace emu

struct emu_graph {
  std::vector<emu::Rec> nodes;
};
static void no_capture(const char* what)
{
  if (emu::capturing()) {
    fprintf(stderr, "EMU: %s recorded during graph capture\n", what);
    abort();
  }
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
  no_capture("hipMemcpy");
  memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { emu::begin_capture(); return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = new emu_graph{emu::end_capture()}; return hipSuccess; }
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) { *e = new emu_graph{g->nodes}; return hipSuccess; }
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) { emu::replay(e->nodes); return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }

static double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emu error"; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n)
{
  *p = malloc(n ? n : 1);
  if (!*p) return hipErrorOutOfMemory;
  memset(*p, 0xA5, n);  // garbage pattern: uninitialised reads show up as NaN-ish values
  return hipSuccess;
}
hipError_t hipFree(void* p) { free(p); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t)
{
  no_capture("hipMemcpyAsync");
  memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { no_capture("hipMemsetAsync"); memset(d, v, n); return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new emu_event{0}; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { no_capture("hipEventRecord"); e->t = now_ms(); return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b)
{
  *ms = (float)(b->t - a->t);
  return hipSuccess;
}
