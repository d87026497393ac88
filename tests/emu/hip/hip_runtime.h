// CPU emulation of the HIP subset that htr-solver_amd/csrc uses, so that the
// library's own kernel source can be compiled with g++ and run on the host
// (test infrastructure: bounds checking under AddressSanitizer and parity of
// the kernels' arithmetic without a GPU).
//
// Execution model.  A launch runs its blocks one after another; the threads of
// a block are real host threads (one per work item), so __syncthreads is a
// block barrier and the wave operations (__shfl*, __ballot) are exchanges
// between the 64 threads of a wave through a wave barrier.  A thread that
// returns from the kernel leaves both barriers; a wave operation that reads a
// lane which has left aborts the run (on the GPU such a read returns an
// undefined value).  __shared__ variables are function-local statics: one
// block runs at a time, so every block sees its own LDS.  "Device" memory is
// host memory (hipMalloc = malloc), so ASan checks every kernel access.
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __constant__
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorMemoryAllocation = 2 };
typedef struct ihipStream_t *hipStream_t;
typedef struct ihipEvent_t *hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };

struct dim3 {
   unsigned x, y, z;
   dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace hipemu {

struct U3 { unsigned x, y, z; };

// barrier whose participant count can drop (exited threads) and be reset
struct Bar {
   std::mutex m;
   std::condition_variable cv;
   int expected = 0, count = 0;
   uint64_t gen = 0;
   void wait()
   {
      std::unique_lock<std::mutex> l(m);
      const uint64_t g = gen;
      if (++count >= expected) { count = 0; gen++; cv.notify_all(); return; }
      cv.wait(l, [&] { return gen != g; });
   }
   void drop()
   {
      std::unique_lock<std::mutex> l(m);
      expected--;
      if (count > 0 && count >= expected) { count = 0; gen++; cv.notify_all(); }
   }
   void reset(int e) { std::lock_guard<std::mutex> l(m); expected = e; count = 0; }
};

struct Wave {
   Bar bar;
   uint64_t slot[64];
   std::atomic<uint64_t> alive{0};
};

constexpr int kMaxWaves = 16;

struct Block {
   Bar bar;
   Wave waves[kMaxWaves];
};

inline thread_local U3 t_tid, t_bid;
inline U3 g_bdim, g_gdim;
inline std::vector<double> g_dyn;   // dynamic LDS of the running launch (blocks run one at a time)
inline thread_local Block *t_block = nullptr;
inline thread_local Wave *t_wave = nullptr;
inline thread_local int t_lane = 0;

[[noreturn]] inline void fail(const char *what)
{
   fprintf(stderr, "hip emulation: %s (block %u,%u,%u thread %u)\n", what, t_bid.x, t_bid.y, t_bid.z, t_tid.x);
   fflush(stderr);
   abort();
}

template<class T> inline uint64_t bits(T v) { static_assert(sizeof(T) <= 8, ""); uint64_t b = 0; memcpy(&b, &v, sizeof(T)); return b; }
template<class T> inline T unbits(uint64_t b) { T v; memcpy(&v, &b, sizeof(T)); return v; }

// every live lane of the wave deposits `mine`; returns the slot of lane src
inline uint64_t exchange(uint64_t mine, int src)
{
   Wave &w = *t_wave;
   w.slot[t_lane] = mine;
   w.bar.wait();
   if (!(w.alive.load() >> src & 1ull)) fail("wave operation reads a lane that has exited");
   const uint64_t r = w.slot[src];
   w.bar.wait();
   return r;
}

inline uint64_t ballot(bool pred)
{
   Wave &w = *t_wave;
   w.slot[t_lane] = pred ? 1 : 0;
   w.bar.wait();
   const uint64_t alive = w.alive.load();
   uint64_t m = 0;
   for (int l = 0; l < 64; l++) if ((alive >> l & 1ull) && w.slot[l]) m |= 1ull << l;
   w.bar.wait();
   return m;
}

template<class K, class... A>
void launch(K kern, dim3 g, dim3 b, size_t shm, A... args)
{
   g_dyn.assign((shm + 7)/8, 0.0);
   const unsigned nthr = b.x*b.y*b.z;
   if (nthr == 0 || nthr > 64u*kMaxWaves) fail("bad block size");
   const uint64_t nblk = (uint64_t)g.x*g.y*g.z;
   if (nblk == 0) return;
   g_gdim = {g.x, g.y, g.z};
   g_bdim = {b.x, b.y, b.z};
   const int nw = (int)((nthr + 63)/64);
   Block *blk = new Block;
   Bar gate;
   gate.reset((int)nthr);
   auto arm = [&]() {
      blk->bar.reset((int)nthr);
      for (int w = 0; w < nw; w++) {
         const int lanes = (int)std::min<unsigned>(64u, nthr - 64u*w);
         blk->waves[w].bar.reset(lanes);
         blk->waves[w].alive.store(lanes == 64 ? ~0ull : ((1ull << lanes) - 1ull));
      }
   };
   arm();
   std::vector<std::thread> th;
   th.reserve(nthr);
   for (unsigned t = 0; t < nthr; t++) {
      th.emplace_back([&, t]() {
         t_tid = {t % b.x, (t/b.x) % b.y, t/(b.x*b.y)};
         t_block = blk;
         t_wave = &blk->waves[t/64];
         t_lane = (int)(t % 64);
         for (uint64_t k = 0; k < nblk; k++) {
            t_bid = {(unsigned)(k % g.x), (unsigned)((k/g.x) % g.y), (unsigned)(k/((uint64_t)g.x*g.y))};
            kern(args...);
            t_wave->alive.fetch_and(~(1ull << t_lane));
            t_wave->bar.drop();
            blk->bar.drop();
            gate.wait();
            if (t == 0) arm();
            gate.wait();
         }
      });
   }
   for (auto &x : th) x.join();
   delete blk;
}

} // namespace hipemu

#define threadIdx (::hipemu::t_tid)
#define blockIdx (::hipemu::t_bid)
#define blockDim (::hipemu::g_bdim)
#define gridDim (::hipemu::g_gdim)

#define hipLaunchKernelGGL(kern, grid, block, shm, stream, ...) \
   ::hipemu::launch(kern, dim3(grid), dim3(block), (size_t)(shm), __VA_ARGS__)
#define HIP_DYNAMIC_SHARED(type, var) type *var = reinterpret_cast<type *>(::hipemu::g_dyn.data());

// the integer and floating min/max of the HIP device library
#define HIPEMU_MINMAX(T) \
   inline T min(T a, T b) { return a < b ? a : b; } \
   inline T max(T a, T b) { return a < b ? b : a; }
HIPEMU_MINMAX(int)
HIPEMU_MINMAX(unsigned)
HIPEMU_MINMAX(long)
HIPEMU_MINMAX(unsigned long)
HIPEMU_MINMAX(long long)
HIPEMU_MINMAX(unsigned long long)
#undef HIPEMU_MINMAX

inline void __syncthreads() { ::hipemu::t_block->bar.wait(); }

template<class T> inline T __shfl(T v, int src, int width = 64)
{
   const int base = ::hipemu::t_lane & ~(width - 1);
   return ::hipemu::unbits<T>(::hipemu::exchange(::hipemu::bits(v), base + (src & (width - 1))));
}
template<class T> inline T __shfl_up(T v, unsigned d, int width = 64)
{
   const int l = ::hipemu::t_lane;
   const int src = ((l & (width - 1)) >= (int)d) ? l - (int)d : l;
   return ::hipemu::unbits<T>(::hipemu::exchange(::hipemu::bits(v), src));
}
template<class T> inline T __shfl_down(T v, unsigned d, int width = 64)
{
   const int l = ::hipemu::t_lane;
   const int src = ((l & (width - 1)) + (int)d < width) ? l + (int)d : l;
   return ::hipemu::unbits<T>(::hipemu::exchange(::hipemu::bits(v), src));
}
template<class T> inline T __shfl_xor(T v, int m, int width = 64)
{
   const int l = ::hipemu::t_lane;
   const int src = (l & ~(width - 1)) | ((l ^ m) & (width - 1));
   return ::hipemu::unbits<T>(::hipemu::exchange(::hipemu::bits(v), src));
}
inline unsigned long long __ballot(int pred) { return ::hipemu::ballot(pred != 0); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline long long __double_as_longlong(double x) { return ::hipemu::unbits<long long>(::hipemu::bits(x)); }
inline double __longlong_as_double(long long x) { return ::hipemu::unbits<double>(::hipemu::bits(x)); }
inline int __float_as_int(float x) { int i; memcpy(&i, &x, sizeof i); return i; }
inline float __int_as_float(int i) { float x; memcpy(&x, &i, sizeof x); return x; }
template<class T> inline T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline double atomicAdd(double *p, double v)
{
   double o = *p, n;
   do { n = o + v; } while (!__atomic_compare_exchange(p, &o, &n, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
   return o;
}
// wave-uniform values only (the library reads a uniform value through it)
template<class T> inline T __builtin_amdgcn_readfirstlane(T v) { return v; }
// the value of v in lane l of the wave (l uniform): an exchange like __shfl
inline int __builtin_amdgcn_readlane(int v, int l)
{
   return ::hipemu::unbits<int>(::hipemu::exchange(::hipemu::bits(v), l & 63));
}
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
inline void __builtin_amdgcn_wave_barrier() { ::hipemu::t_wave->bar.wait(); }
inline void __builtin_amdgcn_sched_barrier(int) {}

// ---- host API -----------------------------------------------------------------
inline hipError_t hipMalloc(void **p, size_t n) { *p = malloc(n ? n : 1); return *p ? hipSuccess : hipErrorMemoryAllocation; }
template<class T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned = 0) { return hipMalloc(p, n); }
template<class T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned f = 0) { return hipHostMalloc((void **)p, n, f); }
inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t = nullptr) { return hipMemcpy(d, s, n, k); }
inline hipError_t hipMemset(void *d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t = nullptr) { return hipMemset(d, v, n); }
inline hipError_t hipMemcpyToSymbol(const void *sym, const void *src, size_t n, size_t off = 0,
                                    hipMemcpyKind = hipMemcpyHostToDevice)
{ memcpy((char *)sym + off, src, n); return hipSuccess; }
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = (void *)sym; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "hip emulation"; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = (hipEvent_t)malloc(1); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { free(e); return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
// launches run to completion in call order, so streams are distinct handles
// (separate library workspaces) and the waits between them hold already
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (hipStream_t)malloc(1); return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = (hipEvent_t)malloc(1); return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.0f; return hipSuccess; }
