// A CPU stand-in for the subset of HIP that libdml uses, so that the
// library's own kernel and host sources compile with g++ and run on the CPU
// (test infrastructure only: the product is the gfx950 build).
//
// Every GPU thread runs as a fiber of its own.  Wave-level intrinsics
// (ballot, shuffles, readlane, wave barrier) and block barriers are the only
// points where fibers meet: a lane that reaches one parks; when every lane of
// its wave has parked (or exited), the lanes parked at the lowest source
// line form the active set of that instruction (the min-PC reconvergence
// rule, which matches the structured control flow of these kernels), get
// their results and run on.  Blocks run one at a time, so __shared__ data is
// a function-local static.  Device memory is host heap memory; built with
// AddressSanitizer, every kernel access is bounds-checked.
#pragma once

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <cstddef>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __constant__
#define __shared__ static
#define __launch_bounds__(...)
#define __forceinline__ inline
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) ((const void *)&(x))

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct float4 {
  float x, y, z, w;
};
struct uint4 {
  unsigned x, y, z, w;
};
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
struct int4 {
  int x, y, z, w;
};
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }

// ---- the SIMT runtime (tests/simt/simt.cpp) -------------------------------
namespace sim {
struct Lane;
extern Lane *g_cur;
extern dim3 g_block_idx, g_block_dim, g_grid_dim;
unsigned lane_tid(const Lane *l);
// wave exchange: every lane contributes `v`; returns what op computes for it
enum Op { OP_BALLOT, OP_SHFL, OP_SHFL_UP, OP_SHFL_XOR, OP_READLANE, OP_READFIRST, OP_WAVEBAR };
uint64_t wave_op(int op, int site, uint64_t v, int arg, int width);
int block_barrier(int site, int pred);
void launch(dim3 grid, dim3 block, void *stream, std::function<void()> body);
unsigned lane_id();
}  // namespace sim

struct sim_tidx {
  unsigned x, y, z;
};
sim_tidx sim_thread_idx();
#define threadIdx (sim_thread_idx())
#define blockIdx (::sim::g_block_idx)
#define blockDim (::sim::g_block_dim)
#define gridDim (::sim::g_grid_dim)

template <class T>
inline uint64_t sim_bits(T v) {
  static_assert(sizeof(T) <= 8, "payload");
  uint64_t u = 0;
  memcpy(&u, &v, sizeof(T));
  return u;
}
template <class T>
inline T sim_from(uint64_t u) {
  T v;
  memcpy(&v, &u, sizeof(T));
  return v;
}

#define __ballot(p) (::sim::wave_op(::sim::OP_BALLOT, __LINE__, (p) ? 1 : 0, 0, 64))
#define __builtin_amdgcn_ballot_w64(p) __ballot(p)
#define __shfl(v, l, ...) sim_shfl_((v), (l), __LINE__, ::sim::OP_SHFL, ##__VA_ARGS__)
#define __shfl_up(v, l, ...) sim_shfl_((v), (l), __LINE__, ::sim::OP_SHFL_UP, ##__VA_ARGS__)
#define __shfl_xor(v, l, ...) sim_shfl_((v), (l), __LINE__, ::sim::OP_SHFL_XOR, ##__VA_ARGS__)
#define __builtin_amdgcn_readlane(v, l) sim_shfl_((v), (l), __LINE__, ::sim::OP_READLANE)
#define __builtin_amdgcn_readfirstlane(v) sim_shfl_((v), 0, __LINE__, ::sim::OP_READFIRST)
#define __builtin_amdgcn_s_waitcnt(x) ((void)(x))
#define __builtin_amdgcn_wave_barrier() ((void)::sim::wave_op(::sim::OP_WAVEBAR, __LINE__, 0, 0, 64))
#define __builtin_amdgcn_fence(...) ((void)0)
#define __syncthreads() ((void)::sim::block_barrier(__LINE__, 0))
#define __syncthreads_or(p) (::sim::block_barrier(__LINE__, (p) ? 1 : 0))

template <class T>
inline T sim_shfl_(T v, int l, int site, int op, int width = 64) {
  return sim_from<T>(::sim::wave_op(op, site, sim_bits<T>(v), l, width));
}

// DPP moves (gfx9 controls row_shr:1-15, row_bcast:15, row_bcast:31): the
// source lane's value where the lane is written and the source exists,
// else 0 (bound_ctrl) or `old`
template <class T>
inline T sim_update_dpp_(T old, T src, int ctrl, int rm, int bm, bool bc, int site) {
  const int l = (int)::sim::lane_id(), row = l >> 4, bank = (l >> 2) & 3;
  int from = -1;
  if (ctrl >= 0 && ctrl <= 0xff) {  // quad_perm
    from = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
  } else if (ctrl >= 0x111 && ctrl <= 0x11f) {
    if ((l & 15) >= ctrl - 0x110) from = l - (ctrl - 0x110);
  } else if (ctrl == 0x142) {
    if (row >= 1) from = row * 16 - 1;
  } else if (ctrl == 0x143) {
    if (row >= 2) from = 31;
  } else {
    abort();
  }
  const T v = sim_shfl_(src, from < 0 ? l : from, site, ::sim::OP_SHFL);
  if (!((rm >> row) & 1) || !((bm >> bank) & 1)) return old;
  if (from < 0) return bc ? T(0) : old;
  return v;
}
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) \
  sim_update_dpp_((old), (src), (ctrl), (rm), (bm), (bc), __LINE__)

inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned m, unsigned acc) {
  const unsigned l = ::sim::lane_id();
  const unsigned mask = l >= 32 ? 0xffffffffu : ((1u << l) - 1u);
  return acc + (unsigned)__builtin_popcount(m & mask);
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned m, unsigned acc) {
  const unsigned l = ::sim::lane_id();
  const unsigned mask = l < 32 ? 0u : (l == 63 ? 0x7fffffffu : ((1u << (l - 32)) - 1u));
  return acc + (unsigned)__builtin_popcount(m & mask);
}

inline bool __builtin_amdgcn_inverse_ballot_w64(unsigned long long m) {
  return (m >> ::sim::lane_id()) & 1ull;
}
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline float __int_as_float(int v) { return sim_from<float>((uint32_t)v); }
inline int __float_as_int(float v) { return (int)(uint32_t)sim_bits<float>(v); }
inline long long __double_as_longlong(double v) { return (long long)sim_bits<double>(v); }
inline double __longlong_as_double(long long v) { return sim_from<double>((uint64_t)v); }
// out of line, both results always computed: inlined, the host compiler
// may call sincos() at one site and cos() alone at another, and the host
// libm's variants can differ in the last bit
void sincospi(double x, double *s, double *c);

template <class A, class B>
inline typename std::common_type<A, B>::type min(A a, B b) {
  return a < b ? a : b;
}
template <class A, class B>
inline typename std::common_type<A, B>::type max(A a, B b) {
  return a < b ? b : a;
}

// atomics: fibers switch only at the intrinsics above, so a read-modify-write
// is indivisible
template <class T, class U>
inline T atomicAdd(T *p, U v) {
  T o = *p;
  *p = (T)(o + (T)v);
  return o;
}
template <class T, class U>
inline T atomicOr(T *p, U v) {
  T o = *p;
  *p = o | (T)v;
  return o;
}
template <class T, class U>
inline T atomicMax(T *p, U v) {
  T o = *p;
  if ((T)v > o) *p = (T)v;
  return o;
}
template <class T, class U>
inline T atomicMin(T *p, U v) {
  T o = *p;
  if ((T)v < o) *p = (T)v;
  return o;
}
template <class T, class U>
inline T atomicCAS(T *p, U cmp, U v) {
  T o = *p;
  if (o == (T)cmp) *p = (T)v;
  return o;
}
template <class T, class U>
inline T atomicExch(T *p, U v) {
  T o = *p;
  *p = (T)v;
  return o;
}
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_WAVEFRONT 0
template <class T>
inline T __hip_atomic_load(const T *p, int, int) {
  return *p;
}

// ---- host API ------------------------------------------------------------
typedef enum hipError_t { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 } hipError_t;
typedef enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
} hipMemcpyKind;
typedef struct sim_stream *hipStream_t;
typedef struct sim_graph *hipGraph_t;
typedef struct sim_graph *hipGraphExec_t;
typedef struct sim_event *hipEvent_t;
typedef struct sim_graph_node *hipGraphNode_t;
enum { hipStreamNonBlocking = 1, hipStreamDefault = 0 };
enum { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };

hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h,
                            hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemset(void *d, int v, size_t n);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipStreamCreateWithFlags(hipStream_t *st, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t st);
hipError_t hipDeviceSynchronize();
hipError_t hipStreamBeginCapture(hipStream_t st, int mode);
hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t st);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
const char *hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipGetDevice(int *dev);
// a __device__ variable is a host object here: its address is the symbol
hipError_t hipGetSymbolAddress(void **p, const void *sym);

namespace sim {
template <class... P, class... A>
void launch_kernel(void (*k)(P...), dim3 g, dim3 b, size_t, hipStream_t st, A &&...a) {
  std::tuple<typename std::decay<P>::type...> args(static_cast<typename std::decay<P>::type>(a)...);
  launch(g, b, (void *)st, [k, args]() { std::apply(k, args); });
}
}  // namespace sim

#define hipLaunchKernelGGL(K, G, B, S, ST, ...) ::sim::launch_kernel(K, dim3(G), dim3(B), S, ST, ##__VA_ARGS__)
