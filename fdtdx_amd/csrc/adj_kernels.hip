// Discrete adjoint of one forward step (the transpose of update_E and update_H at fixed
// materials), i.e. what jax.vjp of forward() computes inside reversible_fdtd's backward loop
// (src/fdtdx/fdtd/fdtd.py:215-251), written as two time-reversed Yee sweeps.
//
// With w_H = c*inv_mu*lam_H and gamma_H = lam_psiH + S_H^T(-w_H):
//   pass A:  mu_E = lam_E - curlH(w_H) + sum_slabs dE^T(a_H * gamma_H),   lam_psiH <- b_H * gamma_H,
//            grad_inv_eps += mu_E * (c * K^{n+1/2}),  K = curlH(H^n) + S_E(b_E psi_E + a_E dH(H^n))
// With w_E = c*inv_eps*mu_E and gamma_E = lam_psiE + S_E^T(w_E):
//   pass B:  lam_H = lam_H + curlE(w_E) + sum_slabs dH^T(a_E * gamma_E),  lam_psiE <- b_E * gamma_E
// using curlH^T = curlE on zero-padded Yee fields. lam_psi is ping-ponged (read lpsi_in, write
// lpsi_out) because the slab transposes read the neighbouring cell's old value. With kappa != 1
// the correction is (1/kappa - 1) d + psi', so each slab's derivative transpose also receives
// (1/kappa - 1) times its scattered w term.
// In conducting cells the updates are F' = ((1 - x) F +- c inv K) / (1 + x): w_H and w_E gain the
// factor 1 / (1 + x) of their cell, the pass-through of lam_H and lam_E the factor
// (1 - x) / (1 + x) (lam_E's in pass C, after pass B has read mu_E), and the inv_eps derivative
// becomes mu_E (c K - c sigma eta0 E^n) / (1 + x)^2; sigma is a constant (initialization.py:287).
#include "engine_common.h"
#include "kernels.h"
#include "stencil.h"

namespace fdtdx {

__device__ __forceinline__ bool cell_of_a(const StepStatic& P, int& i, int& j, int& k, long long& o, uint32_t& m) {
  const long long lb = logical_block(P.nblocks_pad);
  if (lb >= P.nblocks) return false;
  m = P.tile_mask ? P.tile_mask[lb] : kMaskAll;
  i = (int)(lb / P.bpp);
  const int p = (int)(lb - (long long)i * P.bpp) * kBlock + threadIdx.x;
  if (p >= P.plane) return false;
  j = p / P.nz;
  k = p - j * P.nz;
  o = (long long)i * P.plane + p;
  return true;
}

__device__ __forceinline__ long long slab_q(const SlabK& s, int i, int j, int k) {
  int li = i, lj = j, lk = k;
  if (s.axis == 0) li -= s.lo; else if (s.axis == 1) lj -= s.lo; else lk -= s.lo;
  return ((long long)li * s.sy + lj) * s.sz + lk;
}

// Component differentiated by d1 / d2 of a slab on `axis`.
//   curl E psi (H coefficients): x -> (Ez, Ey), y -> (Ex, Ez), z -> (Ey, Ex)
//   curl H psi (E coefficients): x -> (Hz, Hy), y -> (Hx, Hz), z -> (Hy, Hx)
// Adds v1 to comp1_of(axis) and v2 to comp2_of(axis) without an axis-indexed array.
__device__ __forceinline__ void add_comp(int axis, float v1, float v2, float& x, float& y, float& z) {
  x += axis == 1 ? v1 : (axis == 2 ? v2 : 0.f);
  y += axis == 2 ? v1 : (axis == 0 ? v2 : 0.f);
  z += axis == 0 ? v1 : (axis == 1 ? v2 : 0.f);
}

// lam_H component c at cell (ii, jj, kk), with the PMC zeroing Z_H applied: the H update ends in
// Z_H (update.py:851), so its transpose starts from Z_H lam_H.
__device__ __forceinline__ float lam_h(const StepStatic& P, const AdjDyn& A, uint32_t m, bool walls, int c,
                                       long long oo, int ii, int jj, int kk) {
  if (walls && wall_zero(P, m, FDTDX_WALL_PMC, c, ii, jj, kk)) return 0.f;
  return A.lamH[c * P.N + oo];
}

// inverse permeability of component c at cell oo (the scalar 1.0 for MC = 0; the diagonal entry
// of a full tensor)
template <int MC>
__device__ __forceinline__ float inv_mu_at(const StepStatic& P, const AdjDyn& A, int c, long long oo) {
  return MC == 0 ? 1.0f : A.im[(MC == 3 ? c : (MC == 9 ? 4 * c : 0)) * P.N + oo];
}

// w_H = (c * inv_mu) * (Z_H lam_H) / (1 + x_H) of component c at cell (ii, jj, kk) (x_H = 0
// outside the magnetic conductivity box).
template <int MC>
__device__ __forceinline__ float w_h(const StepStatic& P, const AdjDyn& A, uint32_t m, bool walls, bool lossy,
                                     int c, long long oo, int ii, int jj, int kk) {
  if constexpr (MC == 9) return A.W[c * P.N + oo];  // k_adj_wT, from Z_H lam_H
  const float imv = inv_mu_at<MC>(P, A, c, oo);
  float w = MC == 0 ? P.c * lam_h(P, A, m, walls, c, oo, ii, jj, kk)
                    : (P.c * imv) * lam_h(P, A, m, walls, c, oo, ii, jj, kk);
  float x[3];
  if (lossy && loss_x(P, m, 1, ii, jj, kk, imv, imv, imv, x)) w = w / (1.0f + x[c]);
  return w;
}

// w_E = (c * inv_eps) * mu_E / (1 + x_E) of component c at cell (ii, jj, kk).
template <int EC>
__device__ __forceinline__ float w_e(const StepStatic& P, const AdjDyn& A, uint32_t m, bool lossy, int c,
                                     long long oo, int ii, int jj, int kk) {
  if constexpr (EC == 9) return A.W[c * P.N + oo];  // k_adj_wT, from mu_E
  const float iev = A.ie[(EC == 3 ? c : 0) * P.N + oo];
  float w = (P.c * iev) * A.lamE[c * P.N + oo];
  float x[3];
  if (lossy && loss_x(P, m, 0, ii, jj, kk, iev, iev, iev, x)) w = w / (1.0f + x[c]);
  return w;
}

__device__ __forceinline__ int coord_a(int a, int i, int j, int k) { return a == 0 ? i : (a == 1 ? j : k); }

// E-type four-point mean (fdtd/misc.py:132-171): (F[p] + F[p + e_l] + F[p - e_c] + F[p + e_l - e_c]) / 4.
__device__ __forceinline__ float avg4_e(const StepStatic& P, const float* Fc, long long o, int i, int j, int k, int c,
                                        int l) {
  long long dl, dc;
  const bool vl = next_off(P, l, coord_a(l, i, j, k), dl);
  const bool vc = prev_off(P, c, coord_a(c, i, j, k), dc);
  return (((Fc[o] + (vl ? Fc[o + dl] : 0.f)) + (vc ? Fc[o + dc] : 0.f)) + ((vl && vc) ? Fc[o + dl + dc] : 0.f)) /
         4.0f;
}

// Full tensors: the cotangent on the curl K of a half-step F' = F +- B avg(K), B = c * inv, is
// W_q = B_qq g_q + sum_{r != q} avg^T(B_rq g_r): the transpose of the four-point mean carries row
// r's value back from component r's positions, (v[p] + v[p - e_r] + v[p + e_q] + v[p - e_r + e_q]) / 4
// for E-type means and (v[p] + v[p + e_r] + v[p - e_q] + v[p + e_r - e_q]) / 4 for H-type, v = B_rq g_r.
// g is Z_H lam_H (H half-step, field 1) or mu_E (E half-step, field 0, already PEC-zeroed).
__device__ __forceinline__ float g_adj(const StepStatic& P, const AdjDyn& A, uint32_t m, bool walls, int field, int r,
                                       long long oo, int ii, int jj, int kk) {
  if (field == 0) return A.lamE[r * P.N + oo];
  return lam_h(P, A, m, walls, r, oo, ii, jj, kk);
}

__global__ __launch_bounds__(kBlock) void k_adj_wT(const StepStatic P, const AdjDyn A, int field) {
  int i, j, k;
  long long o;
  uint32_t m;
  if (!cell_of_a(P, i, j, k, o, m)) return;
  const bool walls = field == 1 && ((m >> kMaskWall) & ((1u << P.nwall) - 1u)) != 0u;
  const float* inv = field == 0 ? A.ie : A.im;
  const int ijk[3] = {i, j, k};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    float acc = 0.f;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const float* B = inv + (long long)(3 * r + q) * P.N;
      if (r == q) {
        acc = acc + (P.c * B[o]) * g_adj(P, A, m, walls, field, r, o, i, j, k);
        continue;
      }
      // neighbours along r (H type: +e_r, E type: -e_r) and q (H type: -e_q, E type: +e_q)
      long long dr, dq;
      const bool vr = field == 1 ? next_off(P, r, ijk[r], dr) : prev_off(P, r, ijk[r], dr);
      const bool vq = field == 1 ? prev_off(P, q, ijk[q], dq) : next_off(P, q, ijk[q], dq);
      int nr[3] = {i, j, k}, nq[3] = {i, j, k}, nb[3] = {i, j, k};
      const int sr = field == 1 ? 1 : -1, sq = -sr;
      nr[r] = wrap_coord(ijk[r] + sr, r == 0 ? P.nx : (r == 1 ? P.ny : P.nz));
      nq[q] = wrap_coord(ijk[q] + sq, q == 0 ? P.nx : (q == 1 ? P.ny : P.nz));
      nb[r] = nr[r];
      nb[q] = nq[q];
      const float v0 = (P.c * B[o]) * g_adj(P, A, m, walls, field, r, o, i, j, k);
      const float v1 = vr ? (P.c * B[o + dr]) * g_adj(P, A, m, walls, field, r, o + dr, nr[0], nr[1], nr[2]) : 0.f;
      const float v2 = vq ? (P.c * B[o + dq]) * g_adj(P, A, m, walls, field, r, o + dq, nq[0], nq[1], nq[2]) : 0.f;
      const float v3 = (vr && vq) ? (P.c * B[o + dr + dq]) *
                                        g_adj(P, A, m, walls, field, r, o + dr + dq, nb[0], nb[1], nb[2])
                                  : 0.f;
      acc = acc + (((v0 + v1) + v2) + v3) / 4.0f;
    }
    A.W[q * P.N + o] = acc;
  }
}

// K^{n+1/2} = curl H^n + the CPML correction with psi_E advanced from the frozen value, at one cell.
__device__ __forceinline__ void adj_curl_K(const StepStatic& P, const AdjDyn& A, uint32_t m, int i, int j, int k,
                                           long long o, float* K) {
  long long dj, dk, di;
  const bool vj = prev_off(P, 1, j, dj), vk = prev_off(P, 2, k, dk), vi = prev_off(P, 0, i, di);
  const float* Hx = A.H;
  const float* Hy = A.H + P.N;
  const float* Hz = A.H + 2 * P.N;
  const float hx = Hx[o], hy = Hy[o], hz = Hz[o];
  // unconditional reads (the cell itself stands in for a neighbour outside the domain)
  const long long oj = vj ? o + dj : o, ok = vk ? o + dk : o, oi = vi ? o + di : o;
  const float hz_j = Hz[oj], hx_j = Hx[oj], hy_k = Hy[ok], hx_k = Hx[ok], hz_i = Hz[oi], hy_i = Hy[oi];
  const float dyHz = hz - (vj ? hz_j : 0.f);
  const float dzHy = hy - (vk ? hy_k : 0.f);
  const float dzHx = hx - (vk ? hx_k : 0.f);
  const float dxHz = hz - (vi ? hz_i : 0.f);
  const float dxHy = hy - (vi ? hy_i : 0.f);
  const float dyHx = hx - (vj ? hx_j : 0.f);
  K[0] = dyHz - dzHy;
  K[1] = dzHx - dxHz;
  K[2] = dxHy - dyHx;
  // slabs that can touch this tile, in object order
  for (uint32_t mm = m & ((1u << P.nslab) - 1u); mm; mm &= mm - 1u) {
    const int s = __builtin_ctz(mm);
    const SlabK& sl = P.slab[s];
    const int ax = sl.axis;
    const int coord = ax == 0 ? i : (ax == 1 ? j : k);
    if (coord < sl.lo || coord >= sl.hi) continue;
    const int dep = coord - sl.lo;
    const float b = sl.coef[dep], a = sl.coef[sl.L + dep];
    float d1, d2;
    if (ax == 0) { d1 = dxHz; d2 = dxHy; } else if (ax == 1) { d1 = dyHx; d2 = dyHz; } else { d1 = dzHy; d2 = dzHx; }
    const long long q = slab_q(sl, i, j, k);
    const float* p1 = A.psi + sl.psi_off;
    float c1 = b * p1[q] + a * d1;
    float c2 = b * p1[sl.cells + q] + a * d2;
    if (!sl.kappa_one) {  // corr = (1/kappa - 1) d + psi' (PML.py:184-188)
      const float km1 = sl.coef[2 * sl.L + dep];
      c1 = km1 * d1 + c1;
      c2 = km1 * d2 + c2;
    }
    K[0] = K[0] + (ax == 1 ? c2 : (ax == 2 ? -c1 : 0.f));
    K[1] = K[1] + (ax == 0 ? -c1 : (ax == 2 ? c2 : 0.f));
    K[2] = K[2] + (ax == 0 ? c2 : (ax == 1 ? -c1 : 0.f));
  }
}

// adj_curl_K plus, for the fused reverse, Kr = curl H^n + the CPML correction with psi frozen (the
// curl update_E_reverse applies, PML.py:181-182), from the same derivative and psi reads.
__device__ __forceinline__ void adj_curl_KK(const StepStatic& P, const AdjDyn& A, uint32_t m, int i, int j, int k,
                                            long long o, float* K, float* Kr) {
  long long dj, dk, di;
  const bool vj = prev_off(P, 1, j, dj), vk = prev_off(P, 2, k, dk), vi = prev_off(P, 0, i, di);
  const float* Hx = A.H;
  const float* Hy = A.H + P.N;
  const float* Hz = A.H + 2 * P.N;
  const float hx = Hx[o], hy = Hy[o], hz = Hz[o];
  const long long oj = vj ? o + dj : o, ok = vk ? o + dk : o, oi = vi ? o + di : o;
  const float hz_j = Hz[oj], hx_j = Hx[oj], hy_k = Hy[ok], hx_k = Hx[ok], hz_i = Hz[oi], hy_i = Hy[oi];
  const float dyHz = hz - (vj ? hz_j : 0.f);
  const float dzHy = hy - (vk ? hy_k : 0.f);
  const float dzHx = hx - (vk ? hx_k : 0.f);
  const float dxHz = hz - (vi ? hz_i : 0.f);
  const float dxHy = hy - (vi ? hy_i : 0.f);
  const float dyHx = hx - (vj ? hx_j : 0.f);
  K[0] = Kr[0] = dyHz - dzHy;
  K[1] = Kr[1] = dzHx - dxHz;
  K[2] = Kr[2] = dxHy - dyHx;
  for (uint32_t mm = m & ((1u << P.nslab) - 1u); mm; mm &= mm - 1u) {
    const int s = __builtin_ctz(mm);
    const SlabK& sl = P.slab[s];
    const int ax = sl.axis;
    const int coord = ax == 0 ? i : (ax == 1 ? j : k);
    if (coord < sl.lo || coord >= sl.hi) continue;
    const int dep = coord - sl.lo;
    const float b = sl.coef[dep], a = sl.coef[sl.L + dep];
    float d1, d2;
    if (ax == 0) { d1 = dxHz; d2 = dxHy; } else if (ax == 1) { d1 = dyHx; d2 = dyHz; } else { d1 = dzHy; d2 = dzHx; }
    const long long q = slab_q(sl, i, j, k);
    const float* p1 = A.psi + sl.psi_off;
    const float s1 = p1[q], s2 = p1[sl.cells + q];
    float c1 = b * s1 + a * d1;
    float c2 = b * s2 + a * d2;
    float r1 = s1, r2 = s2;
    if (!sl.kappa_one) {  // corr = (1/kappa - 1) d + psi (PML.py:184-188)
      const float km1 = sl.coef[2 * sl.L + dep];
      c1 = km1 * d1 + c1;
      c2 = km1 * d2 + c2;
      r1 = km1 * d1 + s1;
      r2 = km1 * d2 + s2;
    }
    K[0] = K[0] + (ax == 1 ? c2 : (ax == 2 ? -c1 : 0.f));
    K[1] = K[1] + (ax == 0 ? -c1 : (ax == 2 ? c2 : 0.f));
    K[2] = K[2] + (ax == 0 ? c2 : (ax == 1 ? -c1 : 0.f));
    Kr[0] = Kr[0] + (ax == 1 ? r2 : (ax == 2 ? -r1 : 0.f));
    Kr[1] = Kr[1] + (ax == 0 ? -r1 : (ax == 2 ? r2 : 0.f));
    Kr[2] = Kr[2] + (ax == 0 ? r2 : (ax == 1 ? -r1 : 0.f));
  }
}

// The reverse E half-step (k_reverse_E's arithmetic) of one cell from its frozen-psi curl Kr:
// E^n = ((E^{n+1} + sources(inverse)) (1 + x) - (c Kr) inv_eps) / (1 - x), PEC faces zeroed;
// stores E^n and returns it.
template <int EC>
__device__ __forceinline__ void reverse_E_cell(const StepStatic& P, const AdjDyn& A, uint32_t m, int i, int j, int k,
                                               long long o, const float* Kr, float* En) {
  float* Ex = A.Erev;
  float* Ey = A.Erev + P.N;
  float* Ez = A.Erev + 2 * P.N;
  float ex = Ex[o], ey = Ey[o], ez = Ez[o];
  apply_sources<false>(P, m, A.ie, A.im, 0, A.n, true, i, j, k, o, EC, 0, ex, ey, ez);
  const float iex = A.ie[o];
  const float iey = EC == 3 ? A.ie[P.N + o] : iex;
  const float iez = EC == 3 ? A.ie[2 * P.N + o] : iex;
  const float c = P.c;
  float rx, ry, rz;
  float xl[3];
  if (loss_x(P, m, 0, i, j, k, iex, iey, iez, xl)) {  // update.py:600-606
    rx = (ex * (1.0f + xl[0]) - (c * Kr[0]) * iex) / (1.0f - xl[0]);
    ry = (ey * (1.0f + xl[1]) - (c * Kr[1]) * iey) / (1.0f - xl[1]);
    rz = (ez * (1.0f + xl[2]) - (c * Kr[2]) * iez) / (1.0f - xl[2]);
  } else {
    rx = ex - (c * Kr[0]) * iex;
    ry = ey - (c * Kr[1]) * iey;
    rz = ez - (c * Kr[2]) * iez;
  }
  apply_walls(P, m, FDTDX_WALL_PEC, i, j, k, rx, ry, rz);  // update.py:683
  Ex[o] = rx;
  Ey[o] = ry;
  Ez[o] = rz;
  En[0] = rx;
  En[1] = ry;
  En[2] = rz;
}

// K^{n+1/2} of every cell, for the off-diagonal permittivity gradient.
__global__ __launch_bounds__(kBlock) void k_adj_K(const StepStatic P, const AdjDyn A) {
  int i, j, k;
  long long o;
  uint32_t m;
  if (!cell_of_a(P, i, j, k, o, m)) return;
  float K[3];
  adj_curl_K(P, A, m, i, j, k, o, K);
  A.K[o] = K[0];
  A.K[P.N + o] = K[1];
  A.K[2 * P.N + o] = K[2];
}

// REV: pass A also undoes the E half-step of step n (reverse_E_cell), from the same H^n and psi
// reads, after which E holds E^n.
template <int EC, int MC, bool REV>
__global__ __launch_bounds__(kBlock) void k_adj_A(const StepStatic P, const AdjDyn A) {
  int i, j, k;
  long long o;
  uint32_t m;
  if (!cell_of_a(P, i, j, k, o, m)) return;
  const float c = P.c;
  const bool walls = ((m >> kMaskWall) & ((1u << P.nwall) - 1u)) != 0u;
  // K^{n+1/2} from the reconstructed H^n with the CPML psi advanced from the frozen value (read
  // first, so its loads overlap the lam_H reads below)
  float K[3];
  float En[3] = {0.f, 0.f, 0.f};
  if constexpr (REV) {
    float Kr[3];
    adj_curl_KK(P, A, m, i, j, k, o, K, Kr);
    reverse_E_cell<EC>(P, A, m, i, j, k, o, Kr, En);
  } else {
    adj_curl_K(P, A, m, i, j, k, o, K);
  }
  const bool lossH = ((m >> (kMaskLoss + 1)) & 1u) != 0u;
  // w_H at m and its backward neighbours
  const float wx = w_h<MC>(P, A, m, walls, lossH, 0, o, i, j, k);
  const float wy = w_h<MC>(P, A, m, walls, lossH, 1, o, i, j, k);
  const float wz = w_h<MC>(P, A, m, walls, lossH, 2, o, i, j, k);
  long long dj, dk, di;
  const bool vj = prev_off(P, 1, j, dj), vk = prev_off(P, 2, k, dk), vi = prev_off(P, 0, i, di);
  const int jm = wrap_coord(j - 1, P.ny), km = wrap_coord(k - 1, P.nz), im_ = wrap_coord(i - 1, P.nx);
  // the neighbour reads are issued unconditionally (at the cell itself where the neighbour is
  // outside the domain) and selected afterwards, so that all of them are in flight together
  const long long oj = vj ? o + dj : o, ok = vk ? o + dk : o, oi = vi ? o + di : o;
  const float tx_jm = w_h<MC>(P, A, m, walls, lossH, 0, oj, i, jm, k);
  const float tz_jm = w_h<MC>(P, A, m, walls, lossH, 2, oj, i, jm, k);
  const float tx_km = w_h<MC>(P, A, m, walls, lossH, 0, ok, i, j, km);
  const float ty_km = w_h<MC>(P, A, m, walls, lossH, 1, ok, i, j, km);
  const float ty_im = w_h<MC>(P, A, m, walls, lossH, 1, oi, im_, j, k);
  const float tz_im = w_h<MC>(P, A, m, walls, lossH, 2, oi, im_, j, k);
  const float wx_jm = vj ? tx_jm : 0.f, wz_jm = vj ? tz_jm : 0.f;
  const float wx_km = vk ? tx_km : 0.f, wy_km = vk ? ty_km : 0.f;
  const float wy_im = vi ? ty_im : 0.f, wz_im = vi ? tz_im : 0.f;
  float mu[3];
  mu[0] = A.lamE[o] - ((wz - wz_jm) - (wy - wy_km));
  mu[1] = A.lamE[P.N + o] - ((wx - wx_km) - (wz - wz_im));
  mu[2] = A.lamE[2 * P.N + o] - ((wy - wy_im) - (wx - wx_jm));
  // slabs that can touch this tile, in object order
  for (uint32_t mm = m & ((1u << P.nslab) - 1u); mm; mm &= mm - 1u) {
    const int s = __builtin_ctz(mm);
    const SlabK& sl = P.slab[s];
    const int ax = sl.axis;
    const int coord = ax == 0 ? i : (ax == 1 ? j : k);
    const float* aH = sl.coef + 4 * sl.L;
    const float* bH = sl.coef + 3 * sl.L;
    const float* kH = sl.coef + 5 * sl.L;  // 1/kappa - 1: d also feeds the correction directly
    const bool kap = !sl.kappa_one;
    const float* in1 = A.lpsi_in + sl.psi_off + 2 * sl.cells;
    const float* in2 = in1 + sl.cells;
    // w_H components scattered by this slab: i_s = (a+1)%3 (corr1, minus), j_s = (a+2)%3 (corr2)
    float wi_m, wj_m, wi_p, wj_p;  // at m and at m - e_a
    if (ax == 0) { wi_m = wy; wj_m = wz; wi_p = wy_im; wj_p = wz_im; }
    else if (ax == 1) { wi_m = wz; wj_m = wx; wi_p = wz_jm; wj_p = wx_jm; }
    else { wi_m = wx; wj_m = wy; wi_p = wx_km; wj_p = wy_km; }
    float u1m = 0.f, u2m = 0.f, u1p = 0.f, u2p = 0.f;
    if (coord >= sl.lo && coord < sl.hi) {
      const int dep = coord - sl.lo;
      const long long q = slab_q(sl, i, j, k);
      const float g1 = in1[q] + wi_m;
      const float g2 = in2[q] - wj_m;
      u1m = aH[dep] * g1;
      u2m = aH[dep] * g2;
      if (kap) {
        u1m = u1m + kH[dep] * wi_m;
        u2m = u2m + kH[dep] * (-wj_m);
      }
      float* out1 = A.lpsi_out + sl.psi_off + 2 * sl.cells;
      out1[q] = bH[dep] * g1;
      out1[sl.cells + q] = bH[dep] * g2;
    }
    const int cp = coord - 1;
    if (cp >= sl.lo && cp < sl.hi) {
      const int dep = cp - sl.lo;
      const long long q = slab_q(sl, ax == 0 ? i - 1 : i, ax == 1 ? j - 1 : j, ax == 2 ? k - 1 : k);
      u1p = aH[dep] * (in1[q] + wi_p);
      u2p = aH[dep] * (in2[q] - wj_p);
      if (kap) {
        u1p = u1p + kH[dep] * wi_p;
        u2p = u2p + kH[dep] * (-wj_p);
      }
    }
    add_comp(ax, u1p - u1m, u2p - u2m, mu[0], mu[1], mu[2]);
  }
  // the E update ends in the PEC zeroing Z_E (update.py:521): its transpose applies Z_E to mu_E
  if (walls) apply_walls(P, m, FDTDX_WALL_PEC, i, j, k, mu[0], mu[1], mu[2]);
  float g[3];
  float a2[3];
  if (loss_a2(P, m, i, j, k, a2)) {  // d/d inv_eps of ((1 - x) E + c K inv_eps) / (1 + x)
    float xl[3];
    const float ie0 = A.ie[o];
    loss_x(P, m, 0, i, j, k, ie0, EC == 3 ? A.ie[P.N + o] : ie0, EC == 3 ? A.ie[2 * P.N + o] : ie0, xl);
    for (int cc = 0; cc < 3; ++cc) {
      const float d = 1.0f + xl[cc];
      const float en = REV ? En[cc] : A.E[cc * P.N + o];
      g[cc] = mu[cc] * ((c * K[cc] - a2[cc] * en) / (d * d));
    }
  } else {
    for (int cc = 0; cc < 3; ++cc) g[cc] = mu[cc] * (c * K[cc]);
  }
  if constexpr (EC == 9) {  // d(B_rq avg(K_q))/d inv_eps_rq = c avg(K_q -> r) (K of every cell from k_adj_K)
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float kq = q == r ? A.K[q * P.N + o] : avg4_e(P, A.K + (long long)q * P.N, o, i, j, k, q, r);
        A.grad[(3 * r + q) * P.N + o] += (c * mu[r]) * kq;
      }
  } else if (EC == 3) {
    for (int cc = 0; cc < 3; ++cc) A.grad[cc * P.N + o] += g[cc];
  } else {
    A.grad[o] += (g[0] + g[1]) + g[2];
  }
  A.lamE[o] = mu[0];
  A.lamE[P.N + o] = mu[1];
  A.lamE[2 * P.N + o] = mu[2];
}

template <int EC, int MC>
__global__ __launch_bounds__(kBlock) void k_adj_B(const StepStatic P, const AdjDyn A) {
  int i, j, k;
  long long o;
  uint32_t m;
  if (!cell_of_a(P, i, j, k, o, m)) return;
  const float c = P.c;
  const bool lossE = ((m >> kMaskLoss) & 1u) != 0u;
  // w_E at m and its forward neighbours
  const float wx = w_e<EC>(P, A, m, lossE, 0, o, i, j, k);
  const float wy = w_e<EC>(P, A, m, lossE, 1, o, i, j, k);
  const float wz = w_e<EC>(P, A, m, lossE, 2, o, i, j, k);
  long long dj, dk, di;
  const bool jp = next_off(P, 1, j, dj), kp = next_off(P, 2, k, dk), ip = next_off(P, 0, i, di);
  const int jn = wrap_coord(j + 1, P.ny), kn = wrap_coord(k + 1, P.nz), in_ = wrap_coord(i + 1, P.nx);
  // unconditional neighbour reads (the cell itself outside the domain), selected afterwards
  const long long oj = jp ? o + dj : o, ok = kp ? o + dk : o, oi = ip ? o + di : o;
  const float tx_jp = w_e<EC>(P, A, m, lossE, 0, oj, i, jn, k);
  const float tz_jp = w_e<EC>(P, A, m, lossE, 2, oj, i, jn, k);
  const float tx_kp = w_e<EC>(P, A, m, lossE, 0, ok, i, j, kn);
  const float ty_kp = w_e<EC>(P, A, m, lossE, 1, ok, i, j, kn);
  const float ty_ip = w_e<EC>(P, A, m, lossE, 1, oi, in_, j, k);
  const float tz_ip = w_e<EC>(P, A, m, lossE, 2, oi, in_, j, k);
  const float wx_jp = jp ? tx_jp : 0.f, wz_jp = jp ? tz_jp : 0.f;
  const float wx_kp = kp ? tx_kp : 0.f, wy_kp = kp ? ty_kp : 0.f;
  const float wy_ip = ip ? ty_ip : 0.f, wz_ip = ip ? tz_ip : 0.f;
  const bool walls = ((m >> kMaskWall) & ((1u << P.nwall) - 1u)) != 0u;
  float l0 = lam_h(P, A, m, walls, 0, o, i, j, k);
  float l1 = lam_h(P, A, m, walls, 1, o, i, j, k);
  float l2 = lam_h(P, A, m, walls, 2, o, i, j, k);
  float xh[3];
  if (loss_x(P, m, 1, i, j, k, inv_mu_at<MC>(P, A, 0, o), inv_mu_at<MC>(P, A, 1, o), inv_mu_at<MC>(P, A, 2, o),
             xh)) {  // pass-through of H^n: (1 - x) / (1 + x)
    l0 = ((1.0f - xh[0]) / (1.0f + xh[0])) * l0;
    l1 = ((1.0f - xh[1]) / (1.0f + xh[1])) * l1;
    l2 = ((1.0f - xh[2]) / (1.0f + xh[2])) * l2;
  }
  float lh[3];
  lh[0] = l0 + ((wz_jp - wz) - (wy_kp - wy));
  lh[1] = l1 + ((wx_kp - wx) - (wz_ip - wz));
  lh[2] = l2 + ((wy_ip - wy) - (wx_jp - wx));
  // slabs that can touch this tile, in object order
  for (uint32_t mm = m & ((1u << P.nslab) - 1u); mm; mm &= mm - 1u) {
    const int s = __builtin_ctz(mm);
    const SlabK& sl = P.slab[s];
    const int ax = sl.axis;
    const int coord = ax == 0 ? i : (ax == 1 ? j : k);
    const float* bE = sl.coef;
    const float* aE = sl.coef + sl.L;
    const float* kE = sl.coef + 2 * sl.L;  // 1/kappa - 1
    const bool kap = !sl.kappa_one;
    const float* in1 = A.lpsi_in + sl.psi_off;
    const float* in2 = in1 + sl.cells;
    float wi_m, wj_m, wi_p, wj_p;  // at m and at m + e_a
    if (ax == 0) { wi_m = wy; wj_m = wz; wi_p = wy_ip; wj_p = wz_ip; }
    else if (ax == 1) { wi_m = wz; wj_m = wx; wi_p = wz_jp; wj_p = wx_jp; }
    else { wi_m = wx; wj_m = wy; wi_p = wx_kp; wj_p = wy_kp; }
    float u1m = 0.f, u2m = 0.f, u1p = 0.f, u2p = 0.f;
    if (coord >= sl.lo && coord < sl.hi) {
      const int dep = coord - sl.lo;
      const long long q = slab_q(sl, i, j, k);
      const float g1 = in1[q] - wi_m;
      const float g2 = in2[q] + wj_m;
      u1m = aE[dep] * g1;
      u2m = aE[dep] * g2;
      if (kap) {
        u1m = u1m + kE[dep] * (-wi_m);
        u2m = u2m + kE[dep] * wj_m;
      }
      float* out1 = A.lpsi_out + sl.psi_off;
      out1[q] = bE[dep] * g1;
      out1[sl.cells + q] = bE[dep] * g2;
    }
    const int cn = coord + 1;
    if (cn >= sl.lo && cn < sl.hi) {
      const int dep = cn - sl.lo;
      const long long q = slab_q(sl, ax == 0 ? i + 1 : i, ax == 1 ? j + 1 : j, ax == 2 ? k + 1 : k);
      u1p = aE[dep] * (in1[q] - wi_p);
      u2p = aE[dep] * (in2[q] + wj_p);
      if (kap) {
        u1p = u1p + kE[dep] * (-wi_p);
        u2p = u2p + kE[dep] * wj_p;
      }
    }
    add_comp(ax, u1m - u1p, u2m - u2p, lh[0], lh[1], lh[2]);
  }
  A.lamH[o] = lh[0];
  A.lamH[P.N + o] = lh[1];
  A.lamH[2 * P.N + o] = lh[2];
}

// Pass C: lam_E^n = (1 - x) / (1 + x) * mu_E in the electric conductivity box, one thread per box
// cell (pass B has read mu_E at every neighbour by now).
template <int EC>
__global__ __launch_bounds__(kBlock) void k_adj_C(const StepStatic P, const AdjDyn A) {
  const LossK& L = P.loss[0];
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (q >= L.cells) return;
  const long long yz = (long long)L.by * L.bz;
  const int i = L.lo[0] + (int)(q / yz);
  const long long r = q - (long long)(i - L.lo[0]) * yz;
  const int j = L.lo[1] + (int)(r / L.bz);
  const int k = L.lo[2] + (int)(r - (long long)(j - L.lo[1]) * L.bz);
  const long long o = ((long long)i * P.ny + j) * P.nz + k;
  const float ie0 = A.ie[o];
  float x[3];
  loss_x(P, kMaskAll, 0, i, j, k, ie0, EC == 3 ? A.ie[P.N + o] : ie0, EC == 3 ? A.ie[2 * P.N + o] : ie0, x);
  for (int cc = 0; cc < 3; ++cc) A.lamE[cc * P.N + o] = ((1.0f - x[cc]) / (1.0f + x[cc])) * A.lamE[cc * P.N + o];
}

hipError_t launch_adjoint(const StepStatic& P, const AdjDyn& A, int ec, int mc, hipStream_t s) {
  const dim3 g((unsigned)P.nblocks_pad);
  const dim3 gc((unsigned)((P.loss[0].cells + kBlock - 1) / kBlock));
  using AdjKernel = void (*)(const StepStatic, const AdjDyn);
#define FDTDX_ADJ_PAIR(E_)                                                                        \
  if (ec == E_) {                                                                                 \
    ka = mc == 9 ? k_adj_A<E_, 9, false>                                                          \
                 : (mc == 3 ? k_adj_A<E_, 3, false> : (mc == 1 ? k_adj_A<E_, 1, false> : k_adj_A<E_, 0, false>)); \
    kb = mc == 9 ? k_adj_B<E_, 9> : (mc == 3 ? k_adj_B<E_, 3> : (mc == 1 ? k_adj_B<E_, 1> : k_adj_B<E_, 0>)); \
  }
  AdjKernel ka = nullptr, kb = nullptr;
  FDTDX_ADJ_PAIR(1)
  FDTDX_ADJ_PAIR(3)
  FDTDX_ADJ_PAIR(9)
#undef FDTDX_ADJ_PAIR
  if (A.Erev) {  // pass A with the fused reverse E half-step (diagonal materials only)
    if (ec == 9 || mc == 9) return hipErrorInvalidValue;
    if (ec == 3) ka = mc == 3 ? k_adj_A<3, 3, true> : (mc == 1 ? k_adj_A<3, 1, true> : k_adj_A<3, 0, true>);
    else ka = mc == 3 ? k_adj_A<1, 3, true> : (mc == 1 ? k_adj_A<1, 1, true> : k_adj_A<1, 0, true>);
  }
  if (!ka) return hipErrorInvalidValue;
  // full tensors: the transposed-tensor cotangent of the H half-step, and every cell's K for the
  // permittivity gradient, before pass A; that of the E half-step (from mu_E) before pass B
  if (mc == 9) hipLaunchKernelGGL(k_adj_wT, g, dim3(kBlock), 0, s, P, A, 1);
  if (ec == 9) hipLaunchKernelGGL(k_adj_K, g, dim3(kBlock), 0, s, P, A);
  hipLaunchKernelGGL(ka, g, dim3(kBlock), 0, s, P, A);
  if (ec == 9) hipLaunchKernelGGL(k_adj_wT, g, dim3(kBlock), 0, s, P, A, 0);
  hipLaunchKernelGGL(kb, g, dim3(kBlock), 0, s, P, A);
  if (P.loss[0].comps) {
    if (ec == 3) hipLaunchKernelGGL(k_adj_C<3>, gc, dim3(kBlock), 0, s, P, A);
    else hipLaunchKernelGGL(k_adj_C<1>, gc, dim3(kBlock), 0, s, P, A);
  }
  return hipGetLastError();
}

}  // namespace fdtdx
