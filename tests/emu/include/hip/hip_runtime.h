// CPU emulation of the HIP subset used by hops_select.hip and hops_runtime.cpp,
// for tests/test_select_emulation.py: device memory is host memory, copies
// are memmove, a kernel launch runs one block at a time with one std::thread
// per work-item and __syncthreads as a barrier, a cross-lane read goes
// through a barrier-guarded array.  Built with AddressSanitizer and UBSan, it
// checks every access of the device selection kernels.
#pragma once
#include <algorithm>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>
typedef int hipError_t;
enum { hipSuccess = 0 };
typedef void *hipStream_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
#define hipHostMallocDefault 0
inline const char *hipGetErrorString(hipError_t) { return "emu error"; }
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t = 0) { if (n) memmove(d, s, n); return 0; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { if (n) memmove(d, s, n); return 0; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t = 0) { memset(d, v, n); return 0; }
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned = 0) { *p = malloc(n); return *p ? 0 : 1; }
inline hipError_t hipHostFree(void *p) { free(p); return 0; }
struct int2 { int x, y; };
inline int2 make_int2(int a, int b) { return {a, b}; }
struct double2 { double x, y; };
struct emu_u3 { unsigned x = 0, y = 0, z = 0; };
extern thread_local emu_u3 threadIdx, blockIdx;
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __launch_bounds__(...)
extern std::barrier<> *g_emu_bar;
inline void __syncthreads() { g_emu_bar->arrive_and_wait(); }
template <class T> T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
int emu_readlane(int v, int l);
#define __builtin_amdgcn_readlane(v, l) emu_readlane((v), (l))
inline long long __double_as_longlong(double d) { long long r; memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long v) { double r; memcpy(&r, &v, 8); return r; }
using std::max;
using std::min;
extern long g_emu_launches;
// persistent work-item threads: each launch runs its blocks one after the
// other on the first block.x threads of the pool
struct EmuPool {
  static constexpr unsigned kMax = 256;
  std::vector<std::thread> th;
  std::barrier<> go{kMax + 1}, fin{kMax + 1};
  std::function<void(unsigned)> job;
  unsigned active = 0;
  bool quit = false;
  EmuPool() {
    for (unsigned t = 0; t < kMax; ++t)
      th.emplace_back([this, t]() {
        for (;;) {
          go.arrive_and_wait();
          if (quit) return;
          if (t < active) job(t);
          fin.arrive_and_wait();
        }
      });
  }
  ~EmuPool() {
    quit = true;
    go.arrive_and_wait();
    for (auto &x : th) x.join();
  }
  void run(unsigned n, std::function<void(unsigned)> f) {
    if (n > kMax) abort();
    job = std::move(f);
    active = n;
    go.arrive_and_wait();
    fin.arrive_and_wait();
  }
};
inline EmuPool &emu_pool() {
  static EmuPool p;
  return p;
}
template <class K, class... A>
void hipLaunchKernelGGL(K k, dim3 grid, dim3 block, size_t, hipStream_t, A... args) {
  ++g_emu_launches;
  for (unsigned bx = 0; bx < grid.x; ++bx) {
    std::barrier<> bar(block.x);
    g_emu_bar = &bar;
    emu_pool().run(block.x, [&](unsigned t) {
      threadIdx.x = t;
      blockIdx.x = bx;
      k(args...);
    });
  }
}
