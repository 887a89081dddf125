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

// Stencil probes: the minimal E and H half-steps of an isotropic, lossless grid without CPML
// (zero halo), one cell per lane in row-major order, E <- E + (c * curl H) * inv_eps and
// H <- H - c * curl E. They move the sweeps' algorithmic bytes (40 and 36 B per cell) with the
// sweeps' neighbour reads but none of their descriptor, mask or plane-order work, so the gap
// between a probe and the product sweep on a CPML-free grid is the cost of that machinery.
// XCD = true hands each of the 8 XCDs a contiguous range of blocks (as the sweeps do).
template <bool XCD>
__device__ __forceinline__ int probe_block(int nblocks_pad) {
  if (!XCD) return (int)blockIdx.x;
  const int per = nblocks_pad / 8;
  return ((int)blockIdx.x % 8) * per + (int)blockIdx.x / 8;
}

template <bool XCD, bool NT>
__global__ __launch_bounds__(256) void k_stencil_E(const float* __restrict__ H, float* __restrict__ E,
                                                     const float* __restrict__ ie, float c, int nx, int ny, int nz,
                                                     int nblocks_pad) {
  const int b = probe_block<XCD>(nblocks_pad);
  const int plane = ny * nz, N = plane * nx;
  const int o = b * 256 + (int)threadIdx.x;
  if (o >= N) return;
  const int i = o / plane, r = o - i * plane, j = r / nz, k = r - j * nz;
  const float* Hx = H;
  const float* Hy = H + N;
  const float* Hz = H + 2 * N;
  const float hx = Hx[o], hy = Hy[o], hz = Hz[o];
  const float hx_j = j > 0 ? Hx[o - nz] : 0.f, hz_j = j > 0 ? Hz[o - nz] : 0.f;
  const float hx_k = k > 0 ? Hx[o - 1] : 0.f, hy_k = k > 0 ? Hy[o - 1] : 0.f;
  const float hy_i = i > 0 ? Hy[o - plane] : 0.f, hz_i = i > 0 ? Hz[o - plane] : 0.f;
  const float cx = (hz - hz_j) - (hy - hy_k);
  const float cy = (hx - hx_k) - (hz - hz_i);
  const float cz = (hy - hy_i) - (hx - hx_j);
  const float e = ld<NT>(ie + o);
  st<NT>(E + o, ld<NT>(E + o) + (c * cx) * e);
  st<NT>(E + N + o, ld<NT>(E + N + o) + (c * cy) * e);
  st<NT>(E + 2 * N + o, ld<NT>(E + 2 * N + o) + (c * cz) * e);
}

template <bool XCD, bool NT>
__global__ __launch_bounds__(256) void k_stencil_H(const float* __restrict__ E, float* __restrict__ H, float c,
                                                     int nx, int ny, int nz, int nblocks_pad) {
  const int b = probe_block<XCD>(nblocks_pad);
  const int plane = ny * nz, N = plane * nx;
  const int o = b * 256 + (int)threadIdx.x;
  if (o >= N) return;
  const int i = o / plane, r = o - i * plane, j = r / nz, k = r - j * nz;
  const float* Ex = E;
  const float* Ey = E + N;
  const float* Ez = E + 2 * N;
  const float ex = Ex[o], ey = Ey[o], ez = Ez[o];
  const float ex_j = j < ny - 1 ? Ex[o + nz] : 0.f, ez_j = j < ny - 1 ? Ez[o + nz] : 0.f;
  const float ex_k = k < nz - 1 ? Ex[o + 1] : 0.f, ey_k = k < nz - 1 ? Ey[o + 1] : 0.f;
  const float ey_i = i < nx - 1 ? Ey[o + plane] : 0.f, ez_i = i < nx - 1 ? Ez[o + plane] : 0.f;
  const float cx = (ez_j - ez) - (ey_k - ey);
  const float cy = (ex_k - ex) - (ez_i - ez);
  const float cz = (ey_i - ey) - (ex_j - ex);
  st<NT>(H + o, ld<NT>(H + o) - c * cx);
  st<NT>(H + N + o, ld<NT>(H + N + o) - c * cy);
  st<NT>(H + 2 * N + o, ld<NT>(H + 2 * N + o) - c * cz);
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

// One E (which = 0) or H (which = 1) stencil probe launch over an nx x ny x nz grid (fields of
// 3 N floats, inv_eps of N floats; N < 2^31); flags bit 0: XCD-contiguous blocks, bit 1:
// nontemporal access to the arrays read once.
extern "C" int fdtdx_stencil_probe(int32_t which, int32_t flags, float* E, float* H, const float* inv_eps, int32_t nx,
                                   int32_t ny, int32_t nz, void* stream) {
  using namespace fdtdx;
  const long long N = (long long)nx * ny * nz;
  if (nx <= 0 || ny <= 0 || nz <= 0 || N >= (1ll << 31) || !E || !H || (which == 0 && !inv_eps)) return FDTDX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int nb = (int)((N + 255) / 256);
  const int nbp = (nb + 7) / 8 * 8;
  const float c = 0.99f / 1.7320508f;
  const bool x = flags & 1, nt = flags & 2;
  if (which == 0) {
    if (x && nt) k_stencil_E<true, true><<<nbp, 256, 0, s>>>(H, E, inv_eps, c, nx, ny, nz, nbp);
    else if (x) k_stencil_E<true, false><<<nbp, 256, 0, s>>>(H, E, inv_eps, c, nx, ny, nz, nbp);
    else if (nt) k_stencil_E<false, true><<<nbp, 256, 0, s>>>(H, E, inv_eps, c, nx, ny, nz, nbp);
    else k_stencil_E<false, false><<<nbp, 256, 0, s>>>(H, E, inv_eps, c, nx, ny, nz, nbp);
  } else {
    if (x && nt) k_stencil_H<true, true><<<nbp, 256, 0, s>>>(E, H, c, nx, ny, nz, nbp);
    else if (x) k_stencil_H<true, false><<<nbp, 256, 0, s>>>(E, H, c, nx, ny, nz, nbp);
    else if (nt) k_stencil_H<false, true><<<nbp, 256, 0, s>>>(E, H, c, nx, ny, nz, nbp);
    else k_stencil_H<false, false><<<nbp, 256, 0, s>>>(E, H, c, nx, ny, nz, nbp);
  }
  return hipGetLastError() == hipSuccess ? FDTDX_OK : FDTDX_EHIP;
}
