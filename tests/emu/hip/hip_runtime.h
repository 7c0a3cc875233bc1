// Test infrastructure: a host (CPU) stand-in for the few HIP device facilities the
// wavefront-cooperative quadrature-point code uses, so that the same source runs on the CPU with
// one std::thread per lane.  Only tests/emu/emu_qp.cpp includes it (through -I tests/emu); the
// product is never built against it.
//
// Model: a block is one 64-lane wavefront; lanes form groups of emu::group_size() lanes.
//   * cross-lane operations (__shfl_xor, __ballot) exchange values among the lanes of the caller's
//     group through a slot array and two group barriers, so a group whose lanes disagree on
//     control flow deadlocks instead of silently reading stale values;
//   * __shfl with width 64 exchanges over the whole wavefront (all 64 lanes must call it);
//   * __builtin_amdgcn_wave_barrier() is a group barrier: a cross-lane LDS dependency without a
//     group sync between writer and reader becomes a data race here (two runs differ), where the
//     GPU would hide it by executing the group in lockstep.
#pragma once
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)

struct EmuDim3 {
  unsigned x = 0, y = 0, z = 0;
};
extern thread_local EmuDim3 threadIdx, blockIdx, blockDim, gridDim;

namespace emu {
struct Wave {
  int G = 16;
  std::barrier<> *wbar = nullptr;   // the 64 lanes
  std::barrier<> *gbar[64] = {};    // one per group, G lanes each (indexed by group)
  double dslot[64];
  long long islot[64];
};
extern thread_local Wave *wave;
inline int lane() { return (int)(threadIdx.x & 63); }
inline void group_sync() { wave->gbar[lane() / wave->G]->arrive_and_wait(); }
inline void wave_sync() { wave->wbar->arrive_and_wait(); }
}  // namespace emu

inline void __builtin_amdgcn_wave_barrier() { emu::group_sync(); }
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)

template <typename T>
inline T emu_xchg_group(T v, int src) {
  emu::Wave *w = emu::wave;
  const int l = emu::lane();
  if constexpr (sizeof(T) == 8 && !std::is_integral<T>::value) w->dslot[l] = (double)v;
  else w->islot[l] = (long long)v;
  emu::group_sync();
  T r;
  if constexpr (sizeof(T) == 8 && !std::is_integral<T>::value) r = (T)w->dslot[src];
  else r = (T)w->islot[src];
  emu::group_sync();
  return r;
}

inline double __shfl_xor(double v, int mask, int width) {
  if (width != emu::wave->G) { std::fprintf(stderr, "emu: shfl_xor width %d != group %d\n", width, emu::wave->G); std::abort(); }
  const int l = emu::lane(), base = l & ~(width - 1);
  return emu_xchg_group<double>(v, base + ((l - base) ^ mask));
}
inline int __shfl_xor(int v, int mask, int width) {
  if (width != emu::wave->G) { std::fprintf(stderr, "emu: shfl_xor width %d != group %d\n", width, emu::wave->G); std::abort(); }
  const int l = emu::lane(), base = l & ~(width - 1);
  return emu_xchg_group<int>(v, base + ((l - base) ^ mask));
}
// wavefront-wide broadcast (width 64): every lane of the wavefront must call it
inline int __shfl(int v, int src, int width = 64) {
  if (width != 64) { std::fprintf(stderr, "emu: __shfl width %d\n", width); std::abort(); }
  emu::Wave *w = emu::wave;
  w->islot[emu::lane()] = v;
  emu::wave_sync();
  const int r = (int)w->islot[src];
  emu::wave_sync();
  return r;
}
// bits of the caller's group only (the groups of a wavefront may sit in different branches)
inline unsigned long long __ballot(int p) {
  emu::Wave *w = emu::wave;
  const int l = emu::lane(), G = w->G, base = l & ~(G - 1);
  w->islot[l] = p ? 1 : 0;
  emu::group_sync();
  unsigned long long m = 0;
  for (int k = 0; k < G; ++k)
    if (w->islot[base + k]) m |= 1ull << (base + k);
  emu::group_sync();
  return m;
}

inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline int __popcll(long long v) { return __builtin_popcountll((unsigned long long)v); }
inline int __clzll(long long v) { return __builtin_clzll((unsigned long long)v); }
inline unsigned atomicAdd(unsigned *p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
