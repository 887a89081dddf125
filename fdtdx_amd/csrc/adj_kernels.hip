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
  const long long lb = logical_block(P.xorder, P.nblocks_pad);
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

// ---------------------------------------------------------------- per-axis CPML (acax)
// A cell's neighbour along an axis also lies in at most one slab of that axis. The per-axis
// terms are added x, y, z, which is the slab loop's object order under acax, with the loop's
// arithmetic: results are bit-identical to the loop.

// One axis for the transposed slab terms: the cell's slab (psi, b, a, 1/kappa - 1) and the slab
// of its neighbour at coordinate + DIR (psi, a, 1/kappa - 1).
struct AxPair {
  AxCell m;
  int sn;               // side of the neighbour's slab (-1: none)
  bool kapn;
  float n1, n2, an, kn;
};

template <int W, int AX, int DIR>
__device__ __forceinline__ void ax_pair(const StepStatic& P, const float* psi, int i, int j, int k, AxPair& t) {
  ax_cell<W, AX>(P, psi, i, j, k, t.m);
  const StepStatic::AxSlab& A = P.axd[AX][0];
  const StepStatic::AxSlab& B = P.axd[AX][1];
  const int cn = (AX == 0 ? i : (AX == 1 ? j : k)) + DIR;
  t.sn = ax_side_of(A, B, cn);
  t.kapn = false;
  t.n1 = t.n2 = t.an = t.kn = 0.f;
  if (t.sn < 0) return;
  const bool h = t.sn == 1;
  const int L = h ? B.L : A.L, dep = cn - (h ? B.lo : A.lo);
  const long long cells = h ? B.cells : A.cells;
  t.kapn = !(h ? B.kappa_one : A.kappa_one);
  const long long q = ax_psi<AX>(P, h ? B.psi_off : A.psi_off, W, cells, L, dep, i, j, k);
  const float* cf = (h ? B.coef : A.coef) + (W ? 3 * L : 0);
  t.n1 = psi[q];
  t.n2 = psi[q + cells];
  t.an = cf[L + dep];
  if (t.kapn) t.kn = cf[2 * L + dep];
}

// The transposed slab terms of one axis (the body of pass A's / pass B's slab loop): lam_psi of the
// cell advanced into out, and u (at the cell) and u' (at the neighbour) scattered into (x, y, z).
// Pass A adds w_i and subtracts w_j inside g and scatters u' - u; pass B the reverse.
template <bool PA>
__device__ __forceinline__ void ax_transpose(const StepStatic& P, int ax, const AxPair& t, float* out, float wi_m,
                                             float wj_m, float wi_p, float wj_p, float& x, float& y, float& z) {
  const float si_m = PA ? wi_m : -wi_m, sj_m = PA ? -wj_m : wj_m;
  const float si_p = PA ? wi_p : -wi_p, sj_p = PA ? -wj_p : wj_p;
  float u1m = 0.f, u2m = 0.f, u1p = 0.f, u2p = 0.f;
  if (t.m.s >= 0) {
    const float g1 = t.m.p1 + si_m;
    const float g2 = t.m.p2 + sj_m;
    u1m = t.m.a * g1;
    u2m = t.m.a * g2;
    if (t.m.kap) {
      u1m = u1m + t.m.km1 * si_m;
      u2m = u2m + t.m.km1 * sj_m;
    }
    out[t.m.q] = t.m.b * g1;
    out[t.m.q + t.m.cells] = t.m.b * g2;
  }
  if (t.sn >= 0) {
    u1p = t.an * (t.n1 + si_p);
    u2p = t.an * (t.n2 + sj_p);
    if (t.kapn) {
      u1p = u1p + t.kn * si_p;
      u2p = u2p + t.kn * sj_p;
    }
  }
  if (t.m.s >= 0 && t.sn >= 0 && t.m.s != t.sn) {  // two adjacent slabs: one add each, in object order
    const float a1 = PA ? 0.f - u1m : u1m - 0.f, a2 = PA ? 0.f - u2m : u2m - 0.f;
    const float b1 = PA ? u1p - 0.f : 0.f - u1p, b2 = PA ? u2p - 0.f : 0.f - u2p;
    const bool lo_first = P.axs[ax][0] < P.axs[ax][1];  // the min side's slab comes first
    if (t.m.s == 0 ? lo_first : !lo_first) {
      add_comp(ax, a1, a2, x, y, z);
      add_comp(ax, b1, b2, x, y, z);
    } else {
      add_comp(ax, b1, b2, x, y, z);
      add_comp(ax, a1, a2, x, y, z);
    }
  } else if (t.m.s >= 0 || t.sn >= 0) {
    add_comp(ax, PA ? u1p - u1m : u1m - u1p, PA ? u2p - u2m : u2m - u2p, x, y, z);
  }
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

// K^{n+1/2} = curl H^n + the CPML correction with psi_E advanced from the frozen value, at one cell;
// with RK also Kr = curl H^n + the CPML correction with psi frozen (the curl update_E_reverse
// applies, PML.py:181-182), from the same derivative and psi reads. AXS: the slabs' psi_E and
// coefficients come preloaded per axis (ax[0..2]) instead of from the slab loop.
template <bool RK, bool AXS>
__device__ __forceinline__ void adj_curl(const StepStatic& P, const AdjDyn& A, uint32_t m, int i, int j, int k,
                                         long long o, const AxCell* ax, float* K, float* Kr) {
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
  if constexpr (RK) {
    Kr[0] = K[0];
    Kr[1] = K[1];
    Kr[2] = K[2];
  }
  // one slab's correction (corr = (1/kappa - 1) d + psi' with kappa != 1, PML.py:184-188)
  auto slab = [&](int ax_, float b, float a, float s1, float s2, bool kap, float km1) {
    float d1, d2;
    if (ax_ == 0) { d1 = dxHz; d2 = dxHy; } else if (ax_ == 1) { d1 = dyHx; d2 = dyHz; } else { d1 = dzHy; d2 = dzHx; }
    float c1 = b * s1 + a * d1;
    float c2 = b * s2 + a * d2;
    float r1 = s1, r2 = s2;
    if (kap) {
      c1 = km1 * d1 + c1;
      c2 = km1 * d2 + c2;
      r1 = km1 * d1 + s1;
      r2 = km1 * d2 + s2;
    }
    K[0] = K[0] + (ax_ == 1 ? c2 : (ax_ == 2 ? -c1 : 0.f));
    K[1] = K[1] + (ax_ == 0 ? -c1 : (ax_ == 2 ? c2 : 0.f));
    K[2] = K[2] + (ax_ == 0 ? c2 : (ax_ == 1 ? -c1 : 0.f));
    if constexpr (RK) {
      Kr[0] = Kr[0] + (ax_ == 1 ? r2 : (ax_ == 2 ? -r1 : 0.f));
      Kr[1] = Kr[1] + (ax_ == 0 ? -r1 : (ax_ == 2 ? r2 : 0.f));
      Kr[2] = Kr[2] + (ax_ == 0 ? r2 : (ax_ == 1 ? -r1 : 0.f));
    }
  };
  if constexpr (AXS) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (ax[a].s >= 0) slab(a, ax[a].b, ax[a].a, ax[a].p1, ax[a].p2, ax[a].kap, ax[a].km1);
    return;
  }
  // slabs that can touch this tile, in object order
  for (uint32_t mm = m & ((1u << P.nslab) - 1u); mm; mm &= mm - 1u) {
    const int s = __builtin_ctz(mm);
    const SlabK& sl = P.slab[s];
    const int ax_ = sl.axis;
    const int coord = ax_ == 0 ? i : (ax_ == 1 ? j : k);
    if (coord < sl.lo || coord >= sl.hi) continue;
    const int dep = coord - sl.lo;
    const long long q = slab_q(sl, i, j, k);
    const float* p1 = A.psi + sl.psi_off;
    slab(ax_, sl.coef[dep], sl.coef[sl.L + dep], p1[q], p1[sl.cells + q], !sl.kappa_one,
         sl.kappa_one ? 0.f : sl.coef[2 * sl.L + dep]);
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
  adj_curl<false, false>(P, A, m, i, j, k, o, nullptr, K, nullptr);
  A.K[o] = K[0];
  A.K[P.N + o] = K[1];
  A.K[2 * P.N + o] = K[2];
}

// REV: pass A also undoes the E half-step of step n (reverse_E_cell), from the same H^n and psi
// reads, after which E holds E^n.
// AXS: per-axis CPML (acax), the slab loads of all three axes issued first.
template <int EC, int MC, bool REV, bool AXS>
__global__ __launch_bounds__(kBlock) void k_adj_A(const StepStatic P, const AdjDyn A) {
  int i, j, k;
  long long o;
  uint32_t m;
  if (!cell_of_a(P, i, j, k, o, m)) return;
  const float c = P.c;
  const bool walls = ((m >> kMaskWall) & ((1u << P.nwall) - 1u)) != 0u;
  AxCell kax[3];  // psi_E of the cell's slabs, for K
  AxPair lax[3];  // lam_psi_H of the cell's and its lower neighbours' slabs
  if constexpr (AXS) {
    ax_cell<0, 0>(P, A.psi, i, j, k, kax[0]);
    ax_cell<0, 1>(P, A.psi, i, j, k, kax[1]);
    ax_cell<0, 2>(P, A.psi, i, j, k, kax[2]);
    ax_pair<1, 0, -1>(P, A.lpsi_in, i, j, k, lax[0]);
    ax_pair<1, 1, -1>(P, A.lpsi_in, i, j, k, lax[1]);
    ax_pair<1, 2, -1>(P, A.lpsi_in, i, j, k, lax[2]);
  }
  // K^{n+1/2} from the reconstructed H^n with the CPML psi advanced from the frozen value (read
  // first, so its loads overlap the lam_H reads below)
  float K[3];
  float En[3] = {0.f, 0.f, 0.f};
  if constexpr (REV) {
    float Kr[3];
    adj_curl<true, AXS>(P, A, m, i, j, k, o, kax, K, Kr);
    reverse_E_cell<EC>(P, A, m, i, j, k, o, Kr, En);
  } else {
    adj_curl<false, AXS>(P, A, m, i, j, k, o, kax, K, nullptr);
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
  if constexpr (AXS) {
    float* out = A.lpsi_out;
    ax_transpose<true>(P, 0, lax[0], out, wy, wz, wy_im, wz_im, mu[0], mu[1], mu[2]);
    ax_transpose<true>(P, 1, lax[1], out, wz, wx, wz_jm, wx_jm, mu[0], mu[1], mu[2]);
    ax_transpose<true>(P, 2, lax[2], out, wx, wy, wx_km, wy_km, mu[0], mu[1], mu[2]);
  }
  // slabs that can touch this tile, in object order
  for (uint32_t mm = AXS ? 0u : m & ((1u << P.nslab) - 1u); mm; mm &= mm - 1u) {
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

template <int EC, int MC, bool AXS>
__global__ __launch_bounds__(kBlock) void k_adj_B(const StepStatic P, const AdjDyn A) {
  int i, j, k;
  long long o;
  uint32_t m;
  if (!cell_of_a(P, i, j, k, o, m)) return;
  const float c = P.c;
  AxPair lax[3];  // lam_psi_E of the cell's and its upper neighbours' slabs
  if constexpr (AXS) {
    ax_pair<0, 0, 1>(P, A.lpsi_in, i, j, k, lax[0]);
    ax_pair<0, 1, 1>(P, A.lpsi_in, i, j, k, lax[1]);
    ax_pair<0, 2, 1>(P, A.lpsi_in, i, j, k, lax[2]);
  }
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
  if constexpr (AXS) {
    float* out = A.lpsi_out;
    ax_transpose<false>(P, 0, lax[0], out, wy, wz, wy_ip, wz_ip, lh[0], lh[1], lh[2]);
    ax_transpose<false>(P, 1, lax[1], out, wz, wx, wz_jp, wx_jp, lh[0], lh[1], lh[2]);
    ax_transpose<false>(P, 2, lax[2], out, wx, wy, wx_kp, wy_kp, lh[0], lh[1], lh[2]);
  }
  // slabs that can touch this tile, in object order
  for (uint32_t mm = AXS ? 0u : m & ((1u << P.nslab) - 1u); mm; mm &= mm - 1u) {
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
  AdjKernel ka = nullptr, kb = nullptr;
  const bool axs = P.acax != 0;
  if (ec == 9 || mc == 9) {  // full tensors: the slab loop
#define FDTDX_ADJ_PAIR(E_)                                                                                    \
  if (ec == E_) {                                                                                             \
    ka = mc == 9 ? k_adj_A<E_, 9, false, false>                                                               \
                 : (mc == 3 ? k_adj_A<E_, 3, false, false>                                                    \
                            : (mc == 1 ? k_adj_A<E_, 1, false, false> : k_adj_A<E_, 0, false, false>));       \
    kb = mc == 9 ? k_adj_B<E_, 9, false>                                                                      \
                 : (mc == 3 ? k_adj_B<E_, 3, false> : (mc == 1 ? k_adj_B<E_, 1, false> : k_adj_B<E_, 0, false>)); \
  }
    FDTDX_ADJ_PAIR(1)
    FDTDX_ADJ_PAIR(3)
    FDTDX_ADJ_PAIR(9)
#undef FDTDX_ADJ_PAIR
  } else {  // diagonal materials: pass A with the fused reverse E half-step when A.Erev is set
#define FDTDX_ADJ_DIAG(E_, M_)                                                                     \
  if (ec == E_ && mc == M_) {                                                                      \
    if (A.Erev) ka = axs ? k_adj_A<E_, M_, true, true> : k_adj_A<E_, M_, true, false>;            \
    else ka = axs ? k_adj_A<E_, M_, false, true> : k_adj_A<E_, M_, false, false>;                 \
    kb = axs ? k_adj_B<E_, M_, true> : k_adj_B<E_, M_, false>;                                     \
  }
    FDTDX_ADJ_DIAG(1, 0)
    FDTDX_ADJ_DIAG(1, 1)
    FDTDX_ADJ_DIAG(1, 3)
    FDTDX_ADJ_DIAG(3, 0)
    FDTDX_ADJ_DIAG(3, 1)
    FDTDX_ADJ_DIAG(3, 3)
#undef FDTDX_ADJ_DIAG
  }
  if (A.Erev && (ec == 9 || mc == 9)) return hipErrorInvalidValue;  // the fused reverse: diagonal only
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
