// Streaming probes: the HBM ceiling the field sweeps are measured against.
//
// k_stream<R, W, V, NT> reads R arrays and writes W arrays of n floats each (array r at in + r*n,
// array w at out + w*n), V floats per lane per access (1: dword, 4: dwordx4), NT: nontemporal
// loads and stores. Each written value depends on every value read at its index, so no read is
// dead. With R:W = 7:3 the probe moves the bytes of the E sweep of an isotropic scene (E, H,
// inv_eps read; E written), with 6:3 those of the H sweep. One element group per lane, no grid
// stride, as the sweeps launch.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fdtdx_hip.h"

namespace fdtdx {

template <int V>
struct VecT;
template <>
struct VecT<1> {
  using T = float;
};
typedef float f4v __attribute__((ext_vector_type(4)));
template <>
struct VecT<4> {
  using T = f4v;
};

template <bool NT, typename T>
__device__ __forceinline__ T ld(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <bool NT, typename T>
__device__ __forceinline__ void st(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

__device__ __forceinline__ float addv(float a, float b) { return a + b; }
__device__ __forceinline__ f4v addv(f4v a, f4v b) { return a + b; }

template <int R, int W, int V, bool NT>
__global__ __launch_bounds__(256) void k_stream(const float* __restrict__ in, float* __restrict__ out, long long n) {
  using T = typename VecT<V>::T;
  const long long nv = n / V;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nv) return;
  const T* src = reinterpret_cast<const T*>(in);
  T* dst = reinterpret_cast<T*>(out);
  T acc = ld<NT>(src + t);
#pragma unroll
  for (int r = 1; r < R; ++r) acc = addv(acc, ld<NT>(src + r * nv + t));
#pragma unroll
  for (int w = 0; w < W; ++w) {
    st<NT>(dst + w * nv + t, acc);
    acc = addv(acc, acc);
  }
  if constexpr (W == 0) {  // read-only probe: keep the loads live with a store that never happens
    const float* a = reinterpret_cast<const float*>(&acc);
    if (a[0] == -1234.5f) dst[t] = acc;
  }
}

template <int R, int W>
static hipError_t launch_rw(int vec, int nt, const float* in, float* out, long long n, hipStream_t s) {
  const long long groups = n / vec;
  const unsigned grid = (unsigned)((groups + 255) / 256);
  if (vec == 4) {
    if (nt) k_stream<R, W, 4, true><<<grid, 256, 0, s>>>(in, out, n);
    else k_stream<R, W, 4, false><<<grid, 256, 0, s>>>(in, out, n);
  } else {
    if (nt) k_stream<R, W, 1, true><<<grid, 256, 0, s>>>(in, out, n);
    else k_stream<R, W, 1, false><<<grid, 256, 0, s>>>(in, out, n);
  }
  return hipGetLastError();
}

}  // namespace fdtdx

extern "C" int fdtdx_stream_probe(int32_t reads, int32_t writes, int32_t vec, int32_t nontemporal, const float* in,
                                  float* out, long long n, void* stream) {
  using namespace fdtdx;
  if ((vec != 1 && vec != 4) || n <= 0 || n % 4 != 0 || !in || !out) return FDTDX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (reads == 1 && writes == 1) e = launch_rw<1, 1>(vec, nontemporal, in, out, n, s);
  else if (reads == 7 && writes == 3) e = launch_rw<7, 3>(vec, nontemporal, in, out, n, s);
  else if (reads == 6 && writes == 3) e = launch_rw<6, 3>(vec, nontemporal, in, out, n, s);
  else if (reads == 4 && writes == 1) e = launch_rw<4, 1>(vec, nontemporal, in, out, n, s);
  else if (reads == 1 && writes == 0) e = launch_rw<1, 0>(vec, nontemporal, in, out, n, s);
  else return FDTDX_EINVAL;
  return e == hipSuccess ? FDTDX_OK : FDTDX_EHIP;
}
