// Forward Yee half-steps: E sweep (curl H + CPML psi_E + ADE + E-side sources + H_prev snapshot)
// and H sweep (curl E + CPML psi_H + H-side sources + interface recording), plus the reverse
// reconstruction sweeps that reuse the same stencils with frozen psi.
//
// Arithmetic follows the reference's float32 expression order exactly (this file is compiled with
// -ffp-contract=off): curl components are raw differences (core/physics/curl.py:265-279,
// 357-366), CPML corrections are added slab by slab in object order (curl.py:284-308,
// 371-394, PML.py:178-188), then E <- E + (c*curl)*inv_eps (update.py:310) and
// H <- H - (c*curl)*inv_mu (update.py:745). Sources are added afterwards in list order
// (update.py:494-519, 824-849).
#include <algorithm>

#include "engine_common.h"
#include "kernels.h"
#include "record_codec.h"
#include "stencil.h"

namespace fdtdx {

// Cell of this lane in logical tile lb (one tile = kBlock consecutive cells of an x-plane, in the
// forward sweeps' permuted plane order when PERM, else row-major).
template <bool PERM>
__device__ __forceinline__ bool cell_of(const StepStatic& P, long long lb, int& i, int& j, int& k, long long& o) {
  if (lb >= P.nblocks) return false;
  i = (int)(lb / P.bpp);
  const int p = (int)(lb - (long long)i * P.bpp) * kBlock + threadIdx.x;
  if (p >= P.plane) return false;
  if (PERM) {
    perm_cell(P, p, j, k);
    o = (long long)i * P.plane + (long long)j * P.nz + k;
  } else {
    j = p / P.nz;
    k = p - j * P.nz;
    o = (long long)i * P.plane + p;
  }
  return true;
}

template <bool PERM>
__device__ __forceinline__ uint32_t tile_mask(const StepStatic& P, long long lb) {
  const uint32_t* m = PERM ? P.fwd_mask : P.tile_mask;
  return m ? m[lb] : kMaskAll;
}

// Backward-difference derivatives of H at o (zero halo, pad_fields core/misc.py:615-641).
// cpart: the part of a complex-field plan being updated (StepDyn::cpart); the neighbours read
// across a periodic min face then carry the Bloch phase.
template <bool X>
__device__ __forceinline__ Deriv deriv_H(const StepStatic& P, const float* H, long long o, int i, int j, int k,
                                         float& hx, float& hy, float& hz, int cpart = 0) {
  const float* Hx = H;
  const float* Hy = H + P.N;
  const float* Hz = H + 2 * P.N;
  hx = Hx[o]; hy = Hy[o]; hz = Hz[o];
  long long dj, dk, di;
  const bool vj = prev_off<X>(P, 1, j, dj), vk = prev_off<X>(P, 2, k, dk), vi = prev_off<X>(P, 0, i, di);
  float hx_jm = vj ? Hx[o + dj] : 0.f;
  float hz_jm = vj ? Hz[o + dj] : 0.f;
  float hx_km = vk ? Hx[o + dk] : 0.f;
  float hy_km = vk ? Hy[o + dk] : 0.f;
  float hy_im = vi ? Hy[o + di] : 0.f;
  float hz_im = vi ? Hz[o + di] : 0.f;
  if (X && cpart && P.bloch) {
    const long long oc = (long long)cpart * 3 * P.N;
    if (vj && j == 0) {
      hx_jm = bloch_ghost(P, 1, 0, hx_jm, Hx[o + dj + oc], cpart);
      hz_jm = bloch_ghost(P, 1, 0, hz_jm, Hz[o + dj + oc], cpart);
    }
    if (vk && k == 0) {
      hx_km = bloch_ghost(P, 2, 0, hx_km, Hx[o + dk + oc], cpart);
      hy_km = bloch_ghost(P, 2, 0, hy_km, Hy[o + dk + oc], cpart);
    }
    if (vi && i == 0) {
      hy_im = bloch_ghost(P, 0, 0, hy_im, Hy[o + di + oc], cpart);
      hz_im = bloch_ghost(P, 0, 0, hz_im, Hz[o + di + oc], cpart);
    }
  }
  Deriv d;
  d.dy_z = hz - hz_jm; d.dz_y = hy - hy_km; d.dz_x = hx - hx_km;
  d.dx_z = hz - hz_im; d.dx_y = hy - hy_im; d.dy_x = hx - hx_jm;
  d.dx_x = d.dy_y = d.dz_z = 0.f;
  return d;
}

// Forward-difference derivatives of E at o.
template <bool X>
__device__ __forceinline__ Deriv deriv_E(const StepStatic& P, const float* E, long long o, int i, int j, int k,
                                         float& ex, float& ey, float& ez, int cpart = 0) {
  const float* Ex = E;
  const float* Ey = E + P.N;
  const float* Ez = E + 2 * P.N;
  ex = Ex[o]; ey = Ey[o]; ez = Ez[o];
  long long dj, dk, di;
  const bool vj = next_off<X>(P, 1, j, dj), vk = next_off<X>(P, 2, k, dk), vi = next_off<X>(P, 0, i, di);
  float ex_jp = vj ? Ex[o + dj] : 0.f;
  float ez_jp = vj ? Ez[o + dj] : 0.f;
  float ex_kp = vk ? Ex[o + dk] : 0.f;
  float ey_kp = vk ? Ey[o + dk] : 0.f;
  float ey_ip = vi ? Ey[o + di] : 0.f;
  float ez_ip = vi ? Ez[o + di] : 0.f;
  if (X && cpart && P.bloch) {  // across a periodic max face: the Bloch phase
    const long long oc = (long long)cpart * 3 * P.N;
    if (vj && j == P.ny - 1) {
      ex_jp = bloch_ghost(P, 1, 1, ex_jp, Ex[o + dj + oc], cpart);
      ez_jp = bloch_ghost(P, 1, 1, ez_jp, Ez[o + dj + oc], cpart);
    }
    if (vk && k == P.nz - 1) {
      ex_kp = bloch_ghost(P, 2, 1, ex_kp, Ex[o + dk + oc], cpart);
      ey_kp = bloch_ghost(P, 2, 1, ey_kp, Ey[o + dk + oc], cpart);
    }
    if (vi && i == P.nx - 1) {
      ey_ip = bloch_ghost(P, 0, 1, ey_ip, Ey[o + di + oc], cpart);
      ez_ip = bloch_ghost(P, 0, 1, ez_ip, Ez[o + di + oc], cpart);
    }
  }
  Deriv d;
  d.dy_z = ez_jp - ez; d.dz_y = ey_kp - ey; d.dz_x = ex_kp - ex;
  d.dx_z = ez_ip - ez; d.dx_y = ey_ip - ey; d.dy_x = ex_jp - ex;
  d.dx_x = d.dy_y = d.dz_z = 0.f;
  return d;
}

// E half-step of step n for cell (i, j, k) at offset o; m is the mask of its tile.
template <int EC, bool U, bool X>
__device__ __forceinline__ void update_E_cell(const StepStatic& P, const StepDyn& D, int n, uint32_t m, int i, int j,
                                              int k, long long o) {
  float hx, hy, hz;
  const Deriv d = deriv_H<X>(P, D.H, o, i, j, k, hx, hy, hz, D.cpart);
  float cx = d.dy_z - d.dz_y;
  float cy = d.dz_x - d.dx_z;
  float cz = d.dx_y - d.dy_x;
  apply_cpml<U>(P, m, D.psi, D.psi, 0, 1, i, j, k, d, cx, cy, cz);
  const float iex = ldg(D.ie + o);
  const float iey = EC == 3 ? ldg(D.ie + P.N + o) : iex;
  const float iez = EC == 3 ? ldg(D.ie + 2 * P.N + o) : iex;
  float* Ex = D.E;
  float* Ey = D.E + P.N;
  float* Ez = D.E + 2 * P.N;
  const float ex = ldg(Ex + o), ey = ldg(Ey + o), ez = ldg(Ez + o);
  const float c = P.c;
  // conductivity (update.py:300-354): E = ((1 - x) E + c curl inv_eps [+ ADE]) / (1 + x)
  float xl[3];
  const bool lossy = X && loss_x(P, m, 0, i, j, k, iex, iey, iez, xl);
  float nx_, ny_, nz_;
  if (lossy) {
    nx_ = (1.0f - xl[0]) * ex + (c * cx) * iex;
    ny_ = (1.0f - xl[1]) * ey + (c * cy) * iey;
    nz_ = (1.0f - xl[2]) * ez + (c * cz) * iez;
  } else {
    nx_ = ex + (c * cx) * iex;
    ny_ = ey + (c * cy) * iey;
    nz_ = ez + (c * cz) * iez;
  }
  // Lorentz/Drude ADE (update.py:316-347): P_hat = c1 P^n + c2 P^{n-1} + c3 E^n,
  // E += inv_eps * sum_p (P^n - P_hat), P^{n-1} <- P^n, P^n <- P_hat.
#pragma unroll
  for (int b = 0; b < FDTDX_MAX_ADE; ++b) {
    if (b >= P.nade) break;
    if (!((m >> (kMaskAde + b)) & 1u)) continue;
    const AdeK& A = P.ade[b];
    if (!in_box(i, j, k, A.lo, A.hi)) continue;
    const long long q = ((long long)(i - A.lo[0]) * A.by + (j - A.lo[1])) * A.bz + (k - A.lo[2]);
    // three slots: P^n in slot n%3, P^{n-1} in (n+2)%3, P^{n+1} written to (n+1)%3
    const long long slot = (long long)A.poles * 3 * A.cells;
    const float* cur = D.ade + A.state_off + (long long)(n % 3) * slot;
    const float* prv = D.ade + A.state_off + (long long)((n + 2) % 3) * slot;
    float* nxt = D.ade + A.state_off + (long long)((n + 1) % 3) * slot;
    float dl[3] = {0.f, 0.f, 0.f};
    const float eo[3] = {ex, ey, ez};
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      for (int p = 0; p < A.poles; ++p) {
        const int cc = A.ccomp == 3 ? comp : 0;
        const long long ci = ((long long)p * A.ccomp + cc) * A.cells + q;
        const float c1 = A.coef[ci];
        const float c2 = A.coef[(long long)A.poles * A.ccomp * A.cells + ci];
        const float c3 = A.coef[2ll * A.poles * A.ccomp * A.cells + ci];
        const long long si = ((long long)p * 3 + comp) * A.cells + q;
        const float pc = cur[si], pp = prv[si];
        const float ph = (c1 * pc + c2 * pp) + c3 * eo[comp];
        dl[comp] = p == 0 ? pc - ph : dl[comp] + (pc - ph);
        nxt[si] = ph;
      }
    }
    nx_ = nx_ + iex * dl[0];
    ny_ = ny_ + iey * dl[1];
    nz_ = nz_ + iez * dl[2];
  }
  if (lossy) {
    nx_ = nx_ / (1.0f + xl[0]);
    ny_ = ny_ / (1.0f + xl[1]);
    nz_ = nz_ / (1.0f + xl[2]);
  }
  apply_sources<U>(P, m, D.ie, D.im, 0, n, false, i, j, k, o, EC, 0, nx_, ny_, nz_);
  apply_walls(P, m, FDTDX_WALL_PEC, i, j, k, nx_, ny_, nz_);
  stg(Ex + o, nx_);
  stg(Ey + o, ny_);
  stg(Ez + o, nz_);
  if (((m >> kMaskSnap) & 1u) && D.snap && in_box(i, j, k, D.slo, D.shi)) {
    const long long sy = D.shi[1] - D.slo[1], sz = D.shi[2] - D.slo[2];
    const long long sc = (long long)(D.shi[0] - D.slo[0]) * sy * sz;
    const long long q = ((long long)(i - D.slo[0]) * sy + (j - D.slo[1])) * sz + (k - D.slo[2]);
    D.snap[q] = hx;
    D.snap[sc + q] = hy;
    D.snap[2 * sc + q] = hz;
  }
}

// E half-step of step n for the cells of tile lb.
template <int EC, bool U, bool X>
__device__ __forceinline__ void update_E_tile(const StepStatic& P, const StepDyn& D, int n, long long lb) {
  int i, j, k;
  long long o;
  if (!cell_of<true>(P, lb, i, j, k, o)) return;
  uint32_t m = tile_mask<true>(P, lb);
  if (X && D.cpart < 0) m &= ~(0xffu << kMaskSrc);  // real sources drive the real part only
  update_E_cell<EC, U, X>(P, D, n, m, i, j, k, o);
}

// collect_interfaces (update.py:1140-1178, fdtd/misc.py:10-35): the float32 E and H of a cell on
// a slab's interface plane into step n's record slot.
__device__ __forceinline__ void record_iface(const StepStatic& P, uint32_t m, char* record, int i, int j, int k,
                                             float ex, float ey, float ez, float hx, float hy, float hz) {
#pragma unroll
  for (int s = 0; s < FDTDX_MAX_CPML; ++s) {
    if (s >= P.nslab) break;
    if (!((m >> s) & 1u)) continue;
    const SlabK& sl = P.slab[s];
    const int coord = sl.axis == 0 ? i : (sl.axis == 1 ? j : k);
    if (coord != sl.iface) continue;
    const long long pc = (long long)sl.pn0 * sl.pn1;
    const long long q = sl.axis == 0 ? (long long)j * P.nz + k : (sl.axis == 1 ? (long long)i * P.nz + k : (long long)i * P.ny + j);
    float* fe = (float*)(record + sl.rec_off_E);
    float* fh = (float*)(record + sl.rec_off_H);
    fe[q] = ex; fe[pc + q] = ey; fe[2 * pc + q] = ez;
    fh[q] = hx; fh[pc + q] = hy; fh[2 * pc + q] = hz;
  }
}

// H half-step of step n for cell (i, j, k) at offset o (m: its tile's mask); record (step n's
// record slot, or null) receives the CPML interface layers in float32 (records with narrow layers
// are written by k_record).
template <int MC, bool U, bool X>
__device__ __forceinline__ void update_H_cell(const StepStatic& P, const StepDyn& D, int n, char* record, uint32_t m,
                                              int i, int j, int k, long long o) {
  float ex, ey, ez;
  const Deriv d = deriv_E<X>(P, D.E, o, i, j, k, ex, ey, ez, D.cpart);
  float cx = d.dy_z - d.dz_y;
  float cy = d.dz_x - d.dx_z;
  float cz = d.dx_y - d.dy_x;
  apply_cpml<U>(P, m, D.psi, D.psi, 1, 1, i, j, k, d, cx, cy, cz);
  float* Hx = D.H;
  float* Hy = D.H + P.N;
  float* Hz = D.H + 2 * P.N;
  const float c = P.c;
  const float imx = MC ? ldg(D.im + o) : 1.0f;
  const float imy = MC == 3 ? ldg(D.im + P.N + o) : imx;
  const float imz = MC == 3 ? ldg(D.im + 2 * P.N + o) : imx;
  const float hx0 = ldg(Hx + o), hy0 = ldg(Hy + o), hz0 = ldg(Hz + o);
  // magnetic conductivity (update.py:741-750): H = ((1 - x) H - c curl inv_mu) / (1 + x)
  float xl[3];
  float hx, hy, hz;
  if (X && loss_x(P, m, 1, i, j, k, imx, imy, imz, xl)) {
    hx = ((1.0f - xl[0]) * hx0 - (c * cx) * imx) / (1.0f + xl[0]);
    hy = ((1.0f - xl[1]) * hy0 - (c * cy) * imy) / (1.0f + xl[1]);
    hz = ((1.0f - xl[2]) * hz0 - (c * cz) * imz) / (1.0f + xl[2]);
  } else {
    hx = hx0 - (c * cx) * imx;
    hy = hy0 - (c * cy) * imy;
    hz = hz0 - (c * cz) * imz;
  }
  apply_sources<U>(P, m, D.ie, D.im, 1, n, false, i, j, k, o, 1, MC, hx, hy, hz);
  apply_walls(P, m, FDTDX_WALL_PMC, i, j, k, hx, hy, hz);
  stg(Hx + o, hx);
  stg(Hy + o, hy);
  stg(Hz + o, hz);
  if (record && (m & 63u)) record_iface(P, m, record, i, j, k, ex, ey, ez, hx, hy, hz);
}

template <int MC, bool U, bool X>
__device__ __forceinline__ void update_H_tile(const StepStatic& P, const StepDyn& D, int n, char* record,
                                              long long lb) {
  int i, j, k;
  long long o;
  if (!cell_of<true>(P, lb, i, j, k, o)) return;
  uint32_t m = tile_mask<true>(P, lb);
  if (X && D.cpart < 0) m &= ~(0xffu << kMaskSrc);
  update_H_cell<MC, U, X>(P, D, n, record, m, i, j, k, o);
}

// ---------------------------------------------------------------- full-tensor materials
// With 9-component inverse permittivity (or permeability) the update is
// F' = F +- B avg(K), B = c * inv (update.py:356-492, 752-822 with A = I, as the engine allows no
// conductivity or poles there): row r is ((B_r0 T_0 + B_r1 T_1) + B_r2 T_2) with T_r = K_r and the
// other curl components averaged onto component r's Yee position. The averages read K at
// neighbouring cells, so the curl (with its CPML correction) is written to a (3, N) buffer by
// one sweep and the tensor product applied by the next.

// Four-point means of component c of a curl onto component l's position (fdtd/misc.py:132-213,
// uniform grid): E type (F[p] + F[p + e_l] + F[p - e_c] + F[p + e_l - e_c]) / 4, H type
// (F[p] + F[p - e_l] + F[p + e_c] + F[p - e_l + e_c]) / 4; zero (or wrapped) beyond the domain.
// Fc points at component c.
__device__ __forceinline__ int coord_of(int a, int i, int j, int k) { return a == 0 ? i : (a == 1 ? j : k); }

template <bool X, bool HTYPE>
__device__ __forceinline__ float avg4(const StepStatic& P, const float* Fc, long long o, int i, int j, int k, int c,
                                      int l) {
  long long dl, dc;
  const bool vl = HTYPE ? prev_off<X>(P, l, coord_of(l, i, j, k), dl) : next_off<X>(P, l, coord_of(l, i, j, k), dl);
  const bool vc = HTYPE ? next_off<X>(P, c, coord_of(c, i, j, k), dc) : prev_off<X>(P, c, coord_of(c, i, j, k), dc);
  const float a0 = Fc[o];
  const float a1 = vl ? Fc[o + dl] : 0.f;
  const float a2 = vc ? Fc[o + dc] : 0.f;
  const float a3 = (vl && vc) ? Fc[o + dl + dc] : 0.f;
  return (((a0 + a1) + a2) + a3) / 4.0f;
}

// Row r of B avg(K) at cell o, B = c * inv (9 components, row-major).
template <bool X, bool HTYPE>
__device__ __forceinline__ float tensor_row(const StepStatic& P, const float* inv, const float* K, long long o, int i,
                                            int j, int k, int r) {
  float t[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float b = P.c * inv[(long long)(3 * r + q) * P.N + o];
    const float kq = q == r ? K[(long long)q * P.N + o] : avg4<X, HTYPE>(P, K + (long long)q * P.N, o, i, j, k, q, r);
    t[q] = b * kq;
  }
  return (t[0] + t[1]) + t[2];
}

// ---- ADE poles in full-tensor permittivity (update.py:361-423): the polarisation recurrence
// P_hat = c1 P^n + c2 P^{n-1} + coupling(E^n), delta = sum_p (P^n - P_hat), folded into the curl
// as K += delta / c before the tensor sweep averages it. With 9-component coupling tensors (oriented
// poles) row i of the coupling is c_ii E_i + sum_{j != i} (c_ij avg(m_ij E_j) + [c_ij != 0]
// avg(c_ij E_j)) / 2, m = [c != 0], avg the E-type four-point mean of component j onto component
// i's position over the points p, p + e_i, p - e_j, p + e_i - e_j (zero beyond the domain, wrapped
// on periodic axes); coefficients outside the ADE box are zero.

// The coefficient c_ij of pole p (ck: its [cells] array) and E_j at the point (i, j, k) + d_l e_l +
// d_c e_c of a four-point mean; zero beyond the domain or the box.
__device__ __forceinline__ void ade_point(const StepStatic& P, const AdeK& A, const float* ck, const float* Ej,
                                          long long o, int ii, int jj, int kk, int l, int dl, int c, int dc,
                                          float& cv, float& ev) {
  int q3[3] = {ii, jj, kk};
  const int n3[3] = {P.nx, P.ny, P.nz};
  long long off = 0;
  bool ok = true;
  if (dl) {
    long long d;
    ok = ok && next_off<true>(P, l, q3[l], d);
    off += d;
    q3[l] = wrap_coord(q3[l] + 1, n3[l]);
  }
  if (dc) {
    long long d;
    ok = ok && prev_off<true>(P, c, q3[c], d);
    off += d;
    q3[c] = wrap_coord(q3[c] - 1, n3[c]);
  }
  ev = ok ? Ej[o + off] : 0.f;
  cv = 0.f;
  if (ok && in_box(q3[0], q3[1], q3[2], A.lo, A.hi))
    cv = ck[((long long)(q3[0] - A.lo[0]) * A.by + (q3[1] - A.lo[1])) * A.bz + (q3[2] - A.lo[2])];
}

// delta[3] of the ADE box A at cell (i, j, k) (box index q), advancing its polarisation state.
__device__ __forceinline__ void tensor_ade(const StepStatic& P, const StepDyn& D, const AdeK& A, int n, int i, int j,
                                           int k, long long o, long long q, float* dl) {
  const long long slot = (long long)A.poles * 3 * A.cells;
  const float* cur = D.ade + A.state_off + (long long)(n % 3) * slot;
  const float* prv = D.ade + A.state_off + (long long)((n + 2) % 3) * slot;
  float* nxt = D.ade + A.state_off + (long long)((n + 1) % 3) * slot;
  const long long per = (long long)A.poles * A.ccomp * A.cells;
  const float e[3] = {D.E[o], D.E[P.N + o], D.E[2 * P.N + o]};
  for (int r = 0; r < 3; ++r) {
    for (int p = 0; p < A.poles; ++p) {
      const long long ci = ((long long)p * A.ccomp + (A.ccomp == 3 ? r : 0)) * A.cells + q;
      const float c1 = A.coef[ci];
      const float c2 = A.coef[per + ci];
      const float* c3 = A.coef + 2 * per + (long long)p * A.c3comp * A.cells;
      float cpl;
      if (A.c3comp == 9) {
        cpl = c3[(long long)(4 * r) * A.cells + q] * e[r];
        for (int cc = 0; cc < 3; ++cc) {
          if (cc == r) continue;
          const float* ck = c3 + (long long)(3 * r + cc) * A.cells;
          const float* Ej = D.E + (long long)cc * P.N;
          float cv[4], ev[4];
          ade_point(P, A, ck, Ej, o, i, j, k, r, 0, cc, 0, cv[0], ev[0]);
          ade_point(P, A, ck, Ej, o, i, j, k, r, 1, cc, 0, cv[1], ev[1]);
          ade_point(P, A, ck, Ej, o, i, j, k, r, 0, cc, 1, cv[2], ev[2]);
          ade_point(P, A, ck, Ej, o, i, j, k, r, 1, cc, 1, cv[3], ev[3]);
          float a1 = 0.f, a2 = 0.f;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const float mt = cv[t] != 0.f ? 1.f : 0.f;
            a1 = t == 0 ? mt * ev[t] : a1 + mt * ev[t];
            a2 = t == 0 ? cv[t] * ev[t] : a2 + cv[t] * ev[t];
          }
          a1 = a1 / 4.0f;
          a2 = a2 / 4.0f;
          const float cij = cv[0];
          cpl = cpl + 0.5f * (cij * a1 + (cij != 0.f ? 1.f : 0.f) * a2);
        }
      } else {
        cpl = c3[(long long)(A.c3comp == 3 ? r : 0) * A.cells + q] * e[r];
      }
      const long long si = ((long long)p * 3 + r) * A.cells + q;
      const float pc = cur[si], pp = prv[si];
      const float ph = (c1 * pc + c2 * pp) + cpl;
      dl[r] = p == 0 ? pc - ph : dl[r] + (pc - ph);
      nxt[si] = ph;
    }
  }
}

// The curl sweep of a full-tensor half-step: K = curl (+ CPML, psi advanced when update = 1 or
// frozen when 0) of H (E = 0) or of E (E = 1), written to D.K; the E sweep also snapshots H_prev.
template <bool X>
__global__ __launch_bounds__(kBlock) void k_tensor_curl(const StepStatic P, const StepDyn D, int field, int update,
                                                        int perm) {
  int i, j, k;
  long long o;
  const long long lb = logical_block(P.xorder, P.nblocks_pad);
  if (perm ? !cell_of<true>(P, lb, i, j, k, o) : !cell_of<false>(P, lb, i, j, k, o)) return;
  const uint32_t m = perm ? tile_mask<true>(P, lb) : tile_mask<false>(P, lb);
  float a, b, c;
  const Deriv d = field == 0 ? deriv_H<X>(P, D.H, o, i, j, k, a, b, c) : deriv_E<X>(P, D.E, o, i, j, k, a, b, c);
  float cx = d.dy_z - d.dz_y;
  float cy = d.dz_x - d.dx_z;
  float cz = d.dx_y - d.dy_x;
  apply_cpml<false>(P, m, D.psi, D.psi, field, update, i, j, k, d, cx, cy, cz);
  if (field == 0 && update && P.nade) {  // ADE poles: curl + delta / c (update.py:416-423)
#pragma unroll
    for (int b = 0; b < FDTDX_MAX_ADE; ++b) {
      if (b >= P.nade) break;
      if (!((m >> (kMaskAde + b)) & 1u)) continue;
      const AdeK& A = P.ade[b];
      if (!in_box(i, j, k, A.lo, A.hi)) continue;
      const long long q = ((long long)(i - A.lo[0]) * A.by + (j - A.lo[1])) * A.bz + (k - A.lo[2]);
      float dl[3];
      tensor_ade(P, D, A, D.n, i, j, k, o, q, dl);
      cx = cx + dl[0] / P.c;
      cy = cy + dl[1] / P.c;
      cz = cz + dl[2] / P.c;
    }
  }
  D.K[o] = cx;
  D.K[P.N + o] = cy;
  D.K[2 * P.N + o] = cz;
  if (P.tsig[field]) {  // the field before its update, for A avg(F); the reverse run's sources first
    const float* F = field == 0 ? D.E : D.H;
    float v0 = F[o], v1 = F[P.N + o], v2 = F[2 * P.N + o];
    if (!update) {
      if (field == 0) apply_sources<false>(P, m, D.ie, D.im, 0, D.n, true, i, j, k, o, 9, 0, v0, v1, v2);
      else apply_sources<false>(P, m, D.ie, D.im, 1, D.n, true, i, j, k, o, 1, 9, v0, v1, v2);
    }
    D.Fo[o] = v0;
    D.Fo[P.N + o] = v1;
    D.Fo[2 * P.N + o] = v2;
  }
  if (field == 0 && update && ((m >> kMaskSnap) & 1u) && D.snap && in_box(i, j, k, D.slo, D.shi)) {
    const long long sy = D.shi[1] - D.slo[1], sz = D.shi[2] - D.slo[2];
    const long long sc = (long long)(D.shi[0] - D.slo[0]) * sy * sz;
    const long long q = ((long long)(i - D.slo[0]) * sy + (j - D.slo[1])) * sz + (k - D.slo[2]);
    D.snap[q] = a;
    D.snap[sc + q] = b;
    D.snap[2 * sc + q] = c;
  }
}

// Inside the conductivity box of a full-tensor half-step: row r of A avg(Fo) +- B avg(K)
// (update.py:470-488, 611-680, 796-820, 960-1000), ((A_r0 X_0 + A_r1 X_1) + A_r2 X_2) +-
// ((B_r0 Y_0 + B_r1 Y_1) + B_r2 Y_2) with X from the pre-update field Fo and Y from the curl K,
// the off-diagonal terms averaged onto component r's position; M = ab + 18 * rev points at the
// reverse matrices. Returns false outside the box.
template <bool X, bool HTYPE>
__device__ __forceinline__ bool tensor_sigma_rows(const StepStatic& P, const StepDyn& D, int w, int rev, float sgn,
                                                  long long o, int i, int j, int k, float* v) {
  const LossK& L = P.loss[w];
  if (!P.tsig[w] || !in_box(i, j, k, L.lo, L.hi)) return false;
  const long long q = ((long long)(i - L.lo[0]) * L.by + (j - L.lo[1])) * L.bz + (k - L.lo[2]);
  const float* M = D.ab[w] + (long long)(rev ? 18 : 0) * L.cells + q;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float x[3], y[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float* Fc = D.Fo + (long long)c * P.N;
      const float* Kc = D.K + (long long)c * P.N;
      const float xf = c == r ? Fc[o] : avg4<X, HTYPE>(P, Fc, o, i, j, k, c, r);
      const float yk = c == r ? Kc[o] : avg4<X, HTYPE>(P, Kc, o, i, j, k, c, r);
      x[c] = M[(long long)(3 * r + c) * L.cells] * xf;
      y[c] = M[(long long)(9 + 3 * r + c) * L.cells] * yk;
    }
    const float a = (x[0] + x[1]) + x[2];
    const float b = (y[0] + y[1]) + y[2];
    v[r] = sgn > 0.f ? a + b : a - b;
  }
  return true;
}

// E' = E + B avg(K) with sources and PEC faces (forward), or E = (E + sources(inverse)) - B avg(K)
// (update_E_reverse, update.py:611-683); A avg(E) +- B avg(K) in the conductivity box.
template <bool X>
__global__ __launch_bounds__(kBlock) void k_tensor_E(const StepStatic P, const StepDyn D, int reverse) {
  int i, j, k;
  long long o;
  const long long lb = logical_block(P.xorder, P.nblocks_pad);
  if (reverse ? !cell_of<false>(P, lb, i, j, k, o) : !cell_of<true>(P, lb, i, j, k, o)) return;
  const uint32_t m = reverse ? tile_mask<false>(P, lb) : tile_mask<true>(P, lb);
  float v[3] = {D.E[o], D.E[P.N + o], D.E[2 * P.N + o]};
  if (tensor_sigma_rows<X, false>(P, D, 0, reverse, reverse ? -1.f : 1.f, o, i, j, k, v)) {
    // the reverse run's sources are already in Fo
  } else {
    if (reverse) apply_sources<false>(P, m, D.ie, D.im, 0, D.n, true, i, j, k, o, 9, 0, v[0], v[1], v[2]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const float t = tensor_row<X, false>(P, D.ie, D.K, o, i, j, k, r);
      v[r] = reverse ? v[r] - t : v[r] + t;
    }
  }
  if (!reverse) apply_sources<false>(P, m, D.ie, D.im, 0, D.n, false, i, j, k, o, 9, 0, v[0], v[1], v[2]);
  apply_walls(P, m, FDTDX_WALL_PEC, i, j, k, v[0], v[1], v[2]);
  D.E[o] = v[0];
  D.E[P.N + o] = v[1];
  D.E[2 * P.N + o] = v[2];
}

// H' = H - B avg(K) with sources, PMC faces and the float32 interface record (forward), or
// H = (H + sources(inverse)) + B avg(K) (update_H_reverse, update.py:960-1004).
template <bool X>
__global__ __launch_bounds__(kBlock) void k_tensor_H(const StepStatic P, const StepDyn D, int reverse) {
  int i, j, k;
  long long o;
  const long long lb = logical_block(P.xorder, P.nblocks_pad);
  if (reverse ? !cell_of<false>(P, lb, i, j, k, o) : !cell_of<true>(P, lb, i, j, k, o)) return;
  const uint32_t m = reverse ? tile_mask<false>(P, lb) : tile_mask<true>(P, lb);
  float v[3] = {D.H[o], D.H[P.N + o], D.H[2 * P.N + o]};
  if (tensor_sigma_rows<X, true>(P, D, 1, reverse, reverse ? 1.f : -1.f, o, i, j, k, v)) {
  } else {
    if (reverse) apply_sources<false>(P, m, D.ie, D.im, 1, D.n, true, i, j, k, o, 1, 9, v[0], v[1], v[2]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const float t = tensor_row<X, true>(P, D.im, D.K, o, i, j, k, r);
      v[r] = reverse ? v[r] + t : v[r] - t;
    }
  }
  if (!reverse) apply_sources<false>(P, m, D.ie, D.im, 1, D.n, false, i, j, k, o, 1, 9, v[0], v[1], v[2]);
  apply_walls(P, m, FDTDX_WALL_PMC, i, j, k, v[0], v[1], v[2]);
  D.H[o] = v[0];
  D.H[P.N + o] = v[1];
  D.H[2 * P.N + o] = v[2];
  if (!reverse && D.record && (m & 63u)) record_iface(P, m, D.record, i, j, k, D.E[o], D.E[P.N + o],
                                                       D.E[2 * P.N + o], v[0], v[1], v[2]);
}

// X: the plan has periodic axes or conductivity (false compiles both out of the sweep).
template <int EC, bool X>
__global__ __launch_bounds__(kBlock) void k_update_E(const StepStatic P, const StepDyn D) {
  update_E_tile<EC, false, X>(P, D, D.n, logical_block(P.xorder, P.nblocks_pad));
}

template <int MC, bool X>
__global__ __launch_bounds__(kBlock) void k_update_H(const StepStatic P, const StepDyn D) {
  update_H_tile<MC, false, X>(P, D, D.n, D.record, logical_block(P.xorder, P.nblocks_pad));
}

// ---------------------------------------------------------------- quad sweeps
// The forward sweeps with four consecutive z cells per lane (nz a multiple of 4): every field
// array moves in dwordx4 loads and stores, the z neighbours inside the quad come from the lane's
// own registers, and a quad tile covers 4 kBlock cells. Per cell the arithmetic is that of
// update_E_cell / update_H_cell, in the same order, so the results are bit-identical.
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v ld4(const float* p) { return *reinterpret_cast<const f4v*>(p); }
__device__ __forceinline__ void st4(float* p, f4v v) { *reinterpret_cast<f4v*>(p) = v; }

constexpr uint32_t kSlabBits = (1u << FDTDX_MAX_CPML) - 1u;
constexpr uint32_t kSnapBit = 1u << kMaskSnap;

// Quad of this lane in quad tile lb: cells (i, j, k0 .. k0 + 3) at offsets o .. o + 3.
__device__ __forceinline__ bool quad_of(const StepStatic& P, long long lb, int& i, int& j, int& k0, long long& o) {
  if (lb >= P.vnblocks) return false;
  i = (int)(lb / P.vbpp);
  const int p = (int)(lb - (long long)i * P.vbpp) * kBlock + threadIdx.x;
  if (p >= P.ny * (P.nz / 4)) return false;
  perm_quad(P, p, j, k0);
  o = (long long)i * P.plane + (long long)j * P.nz + k0;
  return true;
}

// One-sided derivatives of the four cells of a quad, one vector per (axis, component) pair that
// feeds a CPML slab: curl H uses (dx Hz, dx Hy), (dy Hx, dy Hz), (dz Hy, dz Hx), curl E the same
// pairs of E (the order of pml_pair).
struct QuadDeriv {
  f4v x1, x2, y1, y2, z1, z2;
};

// CPML of the four cells of a quad, slab by slab in object order (per cell the order of
// apply_cpml, with cpml_slab's arithmetic). A slab holding all four cells with 16-byte aligned
// psi moves psi1 and psi2 as dwordx4; a slab holding only some of them (a z slab edge inside the
// quad) updates those cells one by one.
__device__ __forceinline__ void cpml_quad(const StepStatic& P, uint32_t mask, float* psi, int which, int i, int j,
                                          int k0, const QuadDeriv& d, f4v& cx, f4v& cy, f4v& cz) {
  uint32_t m = mask & ((1u << P.nslab) - 1u);
  while (m) {
    const int s = __builtin_ctz(m);
    m &= m - 1u;
    const SlabK& sl = P.slab[s];
    const int c0 = sl.axis == 0 ? i : (sl.axis == 1 ? j : k0);
    const int c1 = sl.axis == 2 ? k0 + 3 : c0;
    if (c1 < sl.lo || c0 >= sl.hi) continue;
    const f4v d1 = sl.axis == 0 ? d.x1 : (sl.axis == 1 ? d.y1 : d.z1);
    const f4v d2 = sl.axis == 0 ? d.x2 : (sl.axis == 1 ? d.y2 : d.z2);
    const float* cf = sl.coef + (which ? 3 * sl.L : 0);
    const long long base = sl.psi_off + (which ? 2 * sl.cells : 0);
    const bool whole = c0 >= sl.lo && c1 < sl.hi;
    const long long off = base + slab_index(sl, i, j, whole ? k0 : (k0 < sl.lo ? sl.lo : k0));
    const bool vec = whole && ((off | sl.cells) & 3) == 0;
    f4v p1, p2;
    if (vec) {
      p1 = ld4(psi + off);
      p2 = ld4(psi + off + sl.cells);
    }
    f4v q1, q2;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int coord = sl.axis == 2 ? k0 + e : c0;
      const bool in = coord >= sl.lo && coord < sl.hi;
      const int dep = in ? coord - sl.lo : 0;
      const long long oe = base + slab_index(sl, i, j, sl.axis == 2 ? (in ? k0 + e : sl.lo) : k0 + e);
      const float b = cf[dep], a = cf[sl.L + dep];
      const float o1 = vec ? p1[e] : (in ? ldg(psi + oe) : 0.f);
      const float o2 = vec ? p2[e] : (in ? ldg(psi + oe + sl.cells) : 0.f);
      const float n1 = b * o1 + a * d1[e];
      const float n2 = b * o2 + a * d2[e];
      if (vec) {
        p1[e] = n1;
        p2[e] = n2;
      } else if (in) {
        stg(psi + oe, n1);
        stg(psi + oe + sl.cells, n2);
      }
      float r1 = n1, r2 = n2;
      if (!sl.kappa_one) {
        const float ikm1 = cf[2 * sl.L + dep];
        r1 = ikm1 * d1[e] + n1;
        r2 = ikm1 * d2[e] + n2;
      }
      // cells outside the slab take no correction: add an exact zero
      q1[e] = in ? r1 : 0.f;
      q2[e] = in ? r2 : 0.f;
    }
    if (vec) {
      st4(psi + off, p1);
      st4(psi + off + sl.cells, p2);
    }
    // scatter (curl[(a+1)%3] += -corr1, curl[(a+2)%3] += corr2) as in the one-cell sweeps
    const f4v zero = {0.f, 0.f, 0.f, 0.f};
    const f4v ax_ = sl.axis == 1 ? q2 : (sl.axis == 2 ? -q1 : zero);
    const f4v ay_ = sl.axis == 0 ? -q1 : (sl.axis == 2 ? q2 : zero);
    const f4v az_ = sl.axis == 0 ? q2 : (sl.axis == 1 ? -q1 : zero);
    cx = cx + ax_;
    cy = cy + ay_;
    cz = cz + az_;
  }
}

// Per-axis CPML of a quad (StepStatic::cax): the quad meets at most one slab per axis, the x slab
// fixed by i, the y slab by j and the z slab by k0 (a quad across a z-slab edge updates the cells
// inside the slab one by one). Every
// psi and coefficient load is issued before any is used (quad_slab_loads, ahead of the field
// loads), then the corrections are added x, y, z, i.e. in slab order (quad_slab_apply) with cpml_slab's
// arithmetic, so the results are bit-identical to cpml_quad. Descriptor fields are read for both
// sides of an axis (wave-uniform) and selected per lane.
struct QuadSlabs {
  bool in[3], hi[3];  // the quad meets the axis's slab; of its max side
  unsigned zm;        // cells of the quad inside its z slab (bit e: cell k0 + e)
  f4v p1[3], p2[3];   // psi1, psi2 of the quad in that slab (zero for cells outside it)
};

// Descriptor fields of the slab of one side of axis AX (wave-uniform), and the float offset of psi1
// of cell (i, j, k0) in it (for the z axis k0 - lo may be negative or reach past the slab when the
// quad straddles a slab edge; only the cells inside are addressed).
template <int W, int AX>
__device__ __forceinline__ long long quad_psi_offset(const StepStatic& P, bool hi, int i, int j, int k0, int& dep,
                                                     long long& cells, const float*& cf, int& L, int& kap,
                                                     int& aligned) {
  const StepStatic::AxSlab& A = P.axd[AX][0];
  const StepStatic::AxSlab& B = P.axd[AX][1];
  const int lo = hi ? B.lo : A.lo;
  L = hi ? B.L : A.L;
  cells = hi ? B.cells : A.cells;
  cf = (hi ? B.coef : A.coef) + (W ? 3 * L : 0);
  kap = hi ? B.kappa_one : A.kappa_one;
  aligned = hi ? B.aligned : A.aligned;
  dep = (AX == 0 ? i : (AX == 1 ? j : k0)) - lo;
  const long long base = (hi ? B.psi_off : A.psi_off) + (W ? 2 * cells : 0);
  if (AX == 0) return base + (long long)dep * P.plane + (long long)j * P.nz + k0;
  if (AX == 1) return base + ((long long)i * L + dep) * P.nz + k0;
  return base + ((long long)i * P.ny + j) * L + dep;
}

template <int W, int AX>
__device__ __forceinline__ void quad_slab_load(const StepStatic& P, const float* psi, int i, int j, int k0,
                                               QuadSlabs& S) {
  const StepStatic::AxSlab& A = P.axd[AX][0];
  const StepStatic::AxSlab& B = P.axd[AX][1];
  if (AX == 2) {  // a quad may straddle a z-slab edge (the plan keeps it off the other slab)
    S.hi[AX] = k0 + 3 >= B.lo && B.lo < B.hi;
    S.in[AX] = S.hi[AX] || k0 < A.hi;
    if (!S.in[AX]) return;
    const int lo = S.hi[AX] ? B.lo : A.lo, hi = S.hi[AX] ? B.hi : A.hi;
    S.zm = 0u;
#pragma unroll
    for (int e = 0; e < 4; ++e) S.zm |= (k0 + e >= lo && k0 + e < hi) ? (1u << e) : 0u;
  } else {
    const int c = AX == 0 ? i : j;
    S.hi[AX] = c >= B.lo && c < B.hi;
    S.in[AX] = S.hi[AX] || (c >= A.lo && c < A.hi);
    if (!S.in[AX]) return;
  }
  int dep, L, kap, aligned;
  long long cells;
  const float* cf;
  const long long q = quad_psi_offset<W, AX>(P, S.hi[AX], i, j, k0, dep, cells, cf, L, kap, aligned);
  if (AX != 2 || (aligned && S.zm == 15u)) {
    S.p1[AX] = ld4(psi + q);
    S.p2[AX] = ld4(psi + q + cells);
  } else {
    f4v a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if ((S.zm >> e) & 1u) {
        a[e] = psi[q + e];
        b[e] = psi[q + e + cells];
      }
    S.p1[AX] = a;
    S.p2[AX] = b;
  }
}

template <int W>
__device__ __forceinline__ void quad_slab_loads(const StepStatic& P, const float* psi, int i, int j, int k0,
                                                QuadSlabs& S) {
  quad_slab_load<W, 0>(P, psi, i, j, k0, S);
  quad_slab_load<W, 1>(P, psi, i, j, k0, S);
  quad_slab_load<W, 2>(P, psi, i, j, k0, S);
}

template <int W, int AX>
__device__ __forceinline__ void quad_slab_axis(const StepStatic& P, const QuadSlabs& S, float* psi, int i, int j,
                                               int k0, const QuadDeriv& d, f4v& cx, f4v& cy, f4v& cz) {
  if (!S.in[AX]) return;
  int dep, L, kap, aligned;
  long long cells;
  const float* cf;
  const long long q = quad_psi_offset<W, AX>(P, S.hi[AX], i, j, k0, dep, cells, cf, L, kap, aligned);
  const bool whole = AX != 2 || S.zm == 15u;
  // depths of the four cells (z: clamped into the slab for the cells outside it)
  int de[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) de[e] = AX != 2 ? dep : (dep + e < 0 ? 0 : (dep + e >= L ? L - 1 : dep + e));
  // z slabs with edges on quad boundaries: the quad's four depths are dep .. dep + 3 with dep and L
  // multiples of 4, so each coefficient vector is one aligned dwordx4
  const bool zvec = AX == 2 && whole && aligned;
  f4v b, a;
  if (zvec) {
    b = ld4(cf + dep);
    a = ld4(cf + L + dep);
  } else if (AX == 2) {
    b = f4v{cf[de[0]], cf[de[1]], cf[de[2]], cf[de[3]]};
    a = f4v{cf[L + de[0]], cf[L + de[1]], cf[L + de[2]], cf[L + de[3]]};
  } else {
    const float bv = cf[dep], av = cf[L + dep];
    b = f4v{bv, bv, bv, bv};
    a = f4v{av, av, av, av};
  }
  const f4v d1 = AX == 0 ? d.x1 : (AX == 1 ? d.y1 : d.z1);
  const f4v d2 = AX == 0 ? d.x2 : (AX == 1 ? d.y2 : d.z2);
  const f4v n1 = b * S.p1[AX] + a * d1;
  const f4v n2 = b * S.p2[AX] + a * d2;
  f4v r1 = n1, r2 = n2;
  if (!kap) {  // 1/kappa - 1
    const float* km = cf + 2 * L;
    f4v k;
    if (zvec) k = ld4(km + dep);
    else k = f4v{km[de[0]], km[de[1]], km[de[2]], km[de[3]]};
    r1 = k * d1 + n1;
    r2 = k * d2 + n2;
  }
  if (whole && (AX != 2 || aligned)) {
    st4(psi + q, n1);
    st4(psi + q + cells, n2);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (!((S.zm >> e) & 1u)) {  // outside the slab: no correction (an exact zero)
        r1[e] = 0.f;
        r2[e] = 0.f;
        continue;
      }
      psi[q + e] = n1[e];
      psi[q + e + cells] = n2[e];
    }
  }
  // curl[(a+1)%3] += -corr1, curl[(a+2)%3] += corr2
  if (AX == 0) {
    cy = cy + (-r1);
    cz = cz + r2;
  } else if (AX == 1) {
    cz = cz + (-r1);
    cx = cx + r2;
  } else {
    cx = cx + (-r1);
    cy = cy + r2;
  }
}

template <int W>
__device__ __forceinline__ void quad_slab_apply(const StepStatic& P, const QuadSlabs& S, float* psi, int i, int j,
                                                int k0, const QuadDeriv& d, f4v& cx, f4v& cy, f4v& cz) {
  quad_slab_axis<W, 0>(P, S, psi, i, j, k0, d, cx, cy, cz);
  quad_slab_axis<W, 1>(P, S, psi, i, j, k0, d, cx, cy, cz);
  quad_slab_axis<W, 2>(P, S, psi, i, j, k0, d, cx, cy, cz);
}

// Quad tiles whose mask holds anything beyond CPML slabs and the detector snapshot (sources, ADE
// boxes, PEC / PMC faces, conductivity) are listed in vslow and run one cell per thread in a
// launch of their own, four blocks per listed tile. Cell of block b: false past the last listed
// tile.
__device__ __forceinline__ bool slow_cell(const StepStatic& P, const int* list, long long nlist, long long b, int& i,
                                          int& j, int& k, long long& o, uint32_t& m) {
  const long long s = b >> 2;
  if (s >= nlist) return false;
  const long long lb = list[s];
  m = P.vec_mask[lb];
  i = (int)(lb / P.vbpp);
  const int c = (int)(b & 3) * kBlock + threadIdx.x;  // cell of the quad tile, four per quad
  const int p = (int)(lb - (long long)i * P.vbpp) * kBlock + (c >> 2);
  if (p >= P.ny * (P.nz / 4)) return false;
  int k0;
  perm_quad(P, p, j, k0);
  k = k0 + (c & 3);
  o = (long long)i * P.plane + (long long)j * P.nz + k;
  return true;
}

// The listed quad tiles of the E (k_update_E4s) and H (k_update_H4s) sweeps, one cell per thread,
// launched beside the quad kernels so that those keep a small register budget.
template <int EC, bool X>
__global__ __launch_bounds__(kBlock) void k_update_E4s(const StepStatic P, const StepDyn D) {
  int i, j, k;
  long long o;
  uint32_t ms;
  if (slow_cell(P, P.vslow, P.vnslow, (long long)blockIdx.x, i, j, k, o, ms))
    update_E_cell<EC, false, X>(P, D, D.n, ms, i, j, k, o);
}

template <int MC, bool X>
__global__ __launch_bounds__(kBlock) void k_update_H4s(const StepStatic P, const StepDyn D) {
  int i, j, k;
  long long o;
  uint32_t ms;
  if (slow_cell(P, P.vslowH, P.vnslowH, (long long)blockIdx.x, i, j, k, o, ms))
    update_H_cell<MC, false, X>(P, D, D.n, D.record, ms, i, j, k, o);
}

// CAX: the per-axis CPML (plans with StepStatic::cax), else the slab loop of cpml_quad.
template <int EC, bool X, bool CAX>
__global__ __launch_bounds__(kBlock) void k_update_E4(const StepStatic P, const StepDyn D) {
  int i, j, k0;
  long long o;
  const long long lb = logical_block(P.xorder, P.vnblocks_pad);
  if (!quad_of(P, lb, i, j, k0, o)) return;
  const uint32_t m = P.vec_mask[lb];
  if (m & P.vlistE) return;  // a listed tile: its cells run in k_update_E4s
  QuadSlabs S;
  S.in[0] = S.in[1] = S.in[2] = false;
  S.hi[0] = S.hi[1] = S.hi[2] = false;
  const bool fast_slabs = CAX && (m & kSlabBits);
  if (fast_slabs) quad_slab_loads<0>(P, D.psi, i, j, k0, S);
  const long long N = P.N;
  const float* Hx = D.H;
  const float* Hy = D.H + N;
  const float* Hz = D.H + 2 * N;
  long long dj, di, dk;
  const bool vj = prev_off<X>(P, 1, j, dj), vi = prev_off<X>(P, 0, i, di), vk = prev_off<X>(P, 2, k0, dk);
  const f4v z4 = {0.f, 0.f, 0.f, 0.f};
  const f4v hx = ld4(Hx + o), hy = ld4(Hy + o), hz = ld4(Hz + o);
  const f4v hxj = vj ? ld4(Hx + o + dj) : z4, hzj = vj ? ld4(Hz + o + dj) : z4;
  const f4v hyi = vi ? ld4(Hy + o + di) : z4, hzi = vi ? ld4(Hz + o + di) : z4;
  const float hxk = vk ? Hx[o + dk] : 0.f, hyk = vk ? Hy[o + dk] : 0.f;
  // k - 1 neighbours of the four cells: the previous cell of the quad, the neighbour quad's last
  const f4v hxm = {hxk, hx[0], hx[1], hx[2]};
  const f4v hym = {hyk, hy[0], hy[1], hy[2]};
  QuadDeriv d;
  d.x1 = hz - hzi;  // dx Hz
  d.x2 = hy - hyi;  // dx Hy
  d.y1 = hx - hxj;  // dy Hx
  d.y2 = hz - hzj;  // dy Hz
  d.z1 = hy - hym;  // dz Hy
  d.z2 = hx - hxm;  // dz Hx
  f4v cx = d.y2 - d.z1;  // dy Hz - dz Hy
  f4v cy = d.z2 - d.x1;  // dz Hx - dx Hz
  f4v cz = d.x2 - d.y1;  // dx Hy - dy Hx
  if (fast_slabs) quad_slab_apply<0>(P, S, D.psi, i, j, k0, d, cx, cy, cz);
  else if (!CAX && (m & kSlabBits)) cpml_quad(P, m, D.psi, 0, i, j, k0, d, cx, cy, cz);
  const f4v iex = ld4(D.ie + o);
  const f4v iey = EC == 3 ? ld4(D.ie + N + o) : iex;
  const f4v iez = EC == 3 ? ld4(D.ie + 2 * N + o) : iex;
  float* Ex = D.E;
  float* Ey = D.E + N;
  float* Ez = D.E + 2 * N;
  const f4v c = {P.c, P.c, P.c, P.c};
  st4(Ex + o, ld4(Ex + o) + (c * cx) * iex);
  st4(Ey + o, ld4(Ey + o) + (c * cy) * iey);
  st4(Ez + o, ld4(Ez + o) + (c * cz) * iez);
  if ((m & kSnapBit) && D.snap) {  // H_prev of the detector boxes, as update_E_cell
    const long long sy = D.shi[1] - D.slo[1], sz = D.shi[2] - D.slo[2];
    const long long sc = (long long)(D.shi[0] - D.slo[0]) * sy * sz;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (!in_box(i, j, k0 + e, D.slo, D.shi)) continue;
      const long long q = ((long long)(i - D.slo[0]) * sy + (j - D.slo[1])) * sz + (k0 + e - D.slo[2]);
      D.snap[q] = hx[e];
      D.snap[sc + q] = hy[e];
      D.snap[2 * sc + q] = hz[e];
    }
  }
}

template <int MC, bool X, bool CAX>
__global__ __launch_bounds__(kBlock) void k_update_H4(const StepStatic P, const StepDyn D) {
  int i, j, k0;
  long long o;
  const long long lb = logical_block(P.xorder, P.vnblocks_pad);
  if (!quad_of(P, lb, i, j, k0, o)) return;
  const uint32_t m = P.vec_mask[lb];
  if (m & P.vlistH) return;  // a listed tile: its cells run in k_update_H4s
  QuadSlabs S;
  S.in[0] = S.in[1] = S.in[2] = false;
  S.hi[0] = S.hi[1] = S.hi[2] = false;
  const bool fast_slabs = CAX && (m & kSlabBits);
  if (fast_slabs) quad_slab_loads<1>(P, D.psi, i, j, k0, S);
  const long long N = P.N;
  const float* Ex = D.E;
  const float* Ey = D.E + N;
  const float* Ez = D.E + 2 * N;
  long long dj, di, dk;
  const bool vj = next_off<X>(P, 1, j, dj), vi = next_off<X>(P, 0, i, di), vk = next_off<X>(P, 2, k0 + 3, dk);
  const f4v z4 = {0.f, 0.f, 0.f, 0.f};
  const f4v ex = ld4(Ex + o), ey = ld4(Ey + o), ez = ld4(Ez + o);
  const f4v exj = vj ? ld4(Ex + o + dj) : z4, ezj = vj ? ld4(Ez + o + dj) : z4;
  const f4v eyi = vi ? ld4(Ey + o + di) : z4, ezi = vi ? ld4(Ez + o + di) : z4;
  const float exk = vk ? Ex[o + 3 + dk] : 0.f, eyk = vk ? Ey[o + 3 + dk] : 0.f;
  // k + 1 neighbours: the next cell of the quad, the neighbour quad's first
  const f4v exp_ = {ex[1], ex[2], ex[3], exk};
  const f4v eyp = {ey[1], ey[2], ey[3], eyk};
  QuadDeriv d;
  d.x1 = ezi - ez;   // dx Ez
  d.x2 = eyi - ey;   // dx Ey
  d.y1 = exj - ex;   // dy Ex
  d.y2 = ezj - ez;   // dy Ez
  d.z1 = eyp - ey;   // dz Ey
  d.z2 = exp_ - ex;  // dz Ex
  f4v cx = d.y2 - d.z1;  // dy Ez - dz Ey
  f4v cy = d.z2 - d.x1;  // dz Ex - dx Ez
  f4v cz = d.x2 - d.y1;  // dx Ey - dy Ex
  if (fast_slabs) quad_slab_apply<1>(P, S, D.psi, i, j, k0, d, cx, cy, cz);
  else if (!CAX && (m & kSlabBits)) cpml_quad(P, m, D.psi, 1, i, j, k0, d, cx, cy, cz);
  float* Hx = D.H;
  float* Hy = D.H + N;
  float* Hz = D.H + 2 * N;
  const f4v one4 = {1.f, 1.f, 1.f, 1.f};
  const f4v imx = MC ? ld4(D.im + o) : one4;
  const f4v imy = MC == 3 ? ld4(D.im + N + o) : imx;
  const f4v imz = MC == 3 ? ld4(D.im + 2 * N + o) : imx;
  const f4v c = {P.c, P.c, P.c, P.c};
  const f4v hx = ld4(Hx + o) - (c * cx) * imx;
  const f4v hy = ld4(Hy + o) - (c * cy) * imy;
  const f4v hz = ld4(Hz + o) - (c * cz) * imz;
  st4(Hx + o, hx);
  st4(Hy + o, hy);
  st4(Hz + o, hz);
  if (D.record && (m & kSlabBits)) {
#pragma unroll 1
    for (int e = 0; e < 4; ++e) record_iface(P, m, D.record, i, j, k0 + e, ex[e], ey[e], ez[e], hx[e], hy[e], hz[e]);
  }
}

// ---------------------------------------------------------------- reverse reconstruction
// update_H_reverse (update.py:856-1007): H <- H + sources(inverse); H <- H + (c*curlE(E))*inv_mu
// with frozen psi_H. update_E_reverse (update.py:526-686): E <- E + sources(inverse);
// E <- E - (c*curlH(H))*inv_eps with frozen psi_E. In conducting cells the old field is first
// multiplied by 1 + x and the result divided by 1 - x.
// AXS: per-axis CPML (acax), the psi loads of all three axes issued ahead of the curl's.
template <int MC, bool AXS>
__global__ __launch_bounds__(kBlock) void k_reverse_H(const StepStatic P, const StepDyn D) {
  int i, j, k;
  long long o;
  const long long lb = logical_block(P.xorder, P.nblocks_pad);
  if (!cell_of<false>(P, lb, i, j, k, o)) return;
  const uint32_t m = tile_mask<false>(P, lb);
  AxCell ax[3];
  if constexpr (AXS) {
    ax_cell<1, 0, false>(P, D.psi, i, j, k, ax[0]);
    ax_cell<1, 1, false>(P, D.psi, i, j, k, ax[1]);
    ax_cell<1, 2, false>(P, D.psi, i, j, k, ax[2]);
  }
  float ex, ey, ez;
  const Deriv d = deriv_E<true>(P, D.E, o, i, j, k, ex, ey, ez);
  float cx = d.dy_z - d.dz_y;
  float cy = d.dz_x - d.dx_z;
  float cz = d.dx_y - d.dy_x;
  if constexpr (AXS) cpml_axes_frozen(ax, d, cx, cy, cz);
  else apply_cpml<false>(P, m, D.psi, D.psi, 1, 0, i, j, k, d, cx, cy, cz);
  float* Hx = D.H;
  float* Hy = D.H + P.N;
  float* Hz = D.H + 2 * P.N;
  float hx = Hx[o], hy = Hy[o], hz = Hz[o];
  apply_sources<false>(P, m, D.ie, D.im, 1, D.n, true, i, j, k, o, 1, MC, hx, hy, hz);
  const float c = P.c;
  const float imx = MC ? D.im[o] : 1.0f;
  const float imy = MC == 3 ? D.im[P.N + o] : imx;
  const float imz = MC == 3 ? D.im[2 * P.N + o] : imx;
  float rx, ry, rz;
  float xl[3];
  if (loss_x(P, m, 1, i, j, k, imx, imy, imz, xl)) {  // update.py:921-926
    rx = (hx * (1.0f + xl[0]) + (c * cx) * imx) / (1.0f - xl[0]);
    ry = (hy * (1.0f + xl[1]) + (c * cy) * imy) / (1.0f - xl[1]);
    rz = (hz * (1.0f + xl[2]) + (c * cz) * imz) / (1.0f - xl[2]);
  } else {
    rx = hx + (c * cx) * imx;
    ry = hy + (c * cy) * imy;
    rz = hz + (c * cz) * imz;
  }
  apply_walls(P, m, FDTDX_WALL_PMC, i, j, k, rx, ry, rz);  // update.py:1004
  Hx[o] = rx;
  Hy[o] = ry;
  Hz[o] = rz;
}

template <int EC>
__global__ __launch_bounds__(kBlock) void k_reverse_E(const StepStatic P, const StepDyn D) {
  int i, j, k;
  long long o;
  const long long lb = logical_block(P.xorder, P.nblocks_pad);
  if (!cell_of<false>(P, lb, i, j, k, o)) return;
  const uint32_t m = tile_mask<false>(P, lb);
  float hx, hy, hz;
  const Deriv d = deriv_H<true>(P, D.H, o, i, j, k, hx, hy, hz);
  float cx = d.dy_z - d.dz_y;
  float cy = d.dz_x - d.dx_z;
  float cz = d.dx_y - d.dy_x;
  apply_cpml<false>(P, m, D.psi, D.psi, 0, 0, i, j, k, d, cx, cy, cz);
  float* Ex = D.E;
  float* Ey = D.E + P.N;
  float* Ez = D.E + 2 * P.N;
  float ex = Ex[o], ey = Ey[o], ez = Ez[o];
  apply_sources<false>(P, m, D.ie, D.im, 0, D.n, true, i, j, k, o, EC, 0, ex, ey, ez);
  const float iex = D.ie[o];
  const float iey = EC == 3 ? D.ie[P.N + o] : iex;
  const float iez = EC == 3 ? D.ie[2 * P.N + o] : iex;
  const float c = P.c;
  float rx, ry, rz;
  float xl[3];
  if (loss_x(P, m, 0, i, j, k, iex, iey, iez, xl)) {  // update.py:600-606
    rx = (ex * (1.0f + xl[0]) - (c * cx) * iex) / (1.0f - xl[0]);
    ry = (ey * (1.0f + xl[1]) - (c * cy) * iey) / (1.0f - xl[1]);
    rz = (ez * (1.0f + xl[2]) - (c * cz) * iez) / (1.0f - xl[2]);
  } else {
    rx = ex - (c * cx) * iex;
    ry = ey - (c * cy) * iey;
    rz = ez - (c * cz) * iez;
  }
  apply_walls(P, m, FDTDX_WALL_PEC, i, j, k, rx, ry, rz);  // update.py:683
  Ex[o] = rx;
  Ey[o] = ry;
  Ez[o] = rz;
}

// add_interfaces (update.py:1181-1222) with Recorder.decompress (recorder.py:121-199): decode step
// n's slot, or interpolate between the slots around it, and write it back into E and H. Where
// interface planes of two slabs cross, the later slab's value is the one restored, as in the
// reference's loop over PML objects.
__global__ __launch_bounds__(kBlock) void k_restore(const StepStatic P, const StepDyn D) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  long long acc = 0;
#pragma unroll
  for (int s = 0; s < FDTDX_MAX_CPML; ++s) {
    if (s >= P.nslab) break;
    const SlabK& sl = P.slab[s];
    const long long pc = (long long)sl.pn0 * sl.pn1;
    if (t < acc + pc) {
      const long long q = t - acc;
      const int u = (int)(q / sl.pn1), v = (int)(q - (long long)u * sl.pn1);
      long long o;
      if (sl.axis == 0) o = (long long)sl.iface * P.plane + (long long)u * P.nz + v;
      else if (sl.axis == 1) o = (long long)u * P.plane + (long long)sl.iface * P.nz + v;
      else o = (long long)u * P.plane + (long long)v * P.nz + sl.iface;
      const int ci = (int)(o / P.plane), cj = (int)((o / P.nz) % P.ny), ck = (int)(o % P.nz);
      for (int s2 = s + 1; s2 < P.nslab; ++s2) {
        const SlabK& l2 = P.slab[s2];
        if ((l2.axis == 0 ? ci : (l2.axis == 1 ? cj : ck)) == l2.iface) return;
      }
      const int nE = P.rec_filter_first ? FDTDX_REC_F32 : sl.fmt_E;
      const int nH = P.rec_filter_first ? FDTDX_REC_F32 : sl.fmt_H;
      for (int cmp = 0; cmp < 3; ++cmp) {
        float e = rec_load(D.record + sl.rec_off_E, sl.fmt_E, cmp * pc + q);
        float h = rec_load(D.record + sl.rec_off_H, sl.fmt_H, cmp * pc + q);
        if (D.rec_interp) {
          e = rec_interp(e, rec_load(D.rec_b + sl.rec_off_E, sl.fmt_E, cmp * pc + q), D.rec_f, nE);
          h = rec_interp(h, rec_load(D.rec_b + sl.rec_off_H, sl.fmt_H, cmp * pc + q), D.rec_f, nH);
        }
        D.E[cmp * P.N + o] = e;
        D.H[cmp * P.N + o] = h;
      }
      return;
    }
    acc += pc;
  }
}

// collect_interfaces + Recorder.compress (update.py:1140-1178, recorder.py:70-119) for records
// with narrow layers: after step n's H sweep, encode the interface layers of E^{n+1} and H^{n+1}
// into step n's slot, one thread per interface cell and slab.
__global__ __launch_bounds__(kBlock) void k_record(const StepStatic P, const StepDyn D) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  long long acc = 0;
  for (int s = 0; s < P.nslab; ++s) {
    const SlabK& sl = P.slab[s];
    const long long pc = (long long)sl.pn0 * sl.pn1;
    if (t < acc + pc) {
      const long long q = t - acc;
      const int u = (int)(q / sl.pn1), v = (int)(q - (long long)u * sl.pn1);
      long long o;
      if (sl.axis == 0) o = (long long)sl.iface * P.plane + (long long)u * P.nz + v;
      else if (sl.axis == 1) o = (long long)u * P.plane + (long long)sl.iface * P.nz + v;
      else o = (long long)u * P.plane + (long long)v * P.nz + sl.iface;
      for (int cmp = 0; cmp < 3; ++cmp) {
        rec_store(D.record + sl.rec_off_E, sl.fmt_E, cmp * pc + q, D.E[cmp * P.N + o]);
        rec_store(D.record + sl.rec_off_H, sl.fmt_H, cmp * pc + q, D.H[cmp * P.N + o]);
      }
      return;
    }
    acc += pc;
  }
}

// ---------------------------------------------------------------- launchers
static inline dim3 grid_of(const StepStatic& P) { return dim3((unsigned)P.nblocks_pad); }

static bool extras(const StepStatic& P) {
  return P.wrap[0] || P.wrap[1] || P.wrap[2] || P.loss[0].comps || P.loss[1].comps;
}

using SweepKernel = void (*)(const StepStatic, const StepDyn);

using TensorCurl = void (*)(const StepStatic, const StepDyn, int, int, int);
using TensorSweep = void (*)(const StepStatic, const StepDyn, int);

// One full-tensor half-step: the curl sweep, then the tensor sweep (field 0: E, 1: H).
// The per-cell matrices of a full-tensor half-step with conductivity (compute_anisotropic_update_
// matrices(_reverse), fdtd/misc.py:69-129): M1 = I + f inv sigma, M2 = I - f inv sigma,
// f = c eta / 2 (eta = eta0 for E, 1 / eta0 for H), A = M1^-1 M2, B = c M1^-1 inv,
// A_rev = M2^-1 M1, B_rev = c M2^-1 inv, evaluated in double and rounded once.
__device__ __forceinline__ void inverse3(const double* M, double* R) {
  const double c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8], c02 = M[3] * M[7] - M[4] * M[6];
  const double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
  const double id = 1.0 / det;
  R[0] = c00 * id; R[1] = (M[2] * M[7] - M[1] * M[8]) * id; R[2] = (M[1] * M[5] - M[2] * M[4]) * id;
  R[3] = c01 * id; R[4] = (M[0] * M[8] - M[2] * M[6]) * id; R[5] = (M[2] * M[3] - M[0] * M[5]) * id;
  R[6] = c02 * id; R[7] = (M[1] * M[6] - M[0] * M[7]) * id; R[8] = (M[0] * M[4] - M[1] * M[3]) * id;
}

__device__ __forceinline__ void mul3(const double* X, const double* Y, double* Z) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Z[3 * r + c] = X[3 * r] * Y[c] + X[3 * r + 1] * Y[3 + c] + X[3 * r + 2] * Y[6 + c];
}

__global__ __launch_bounds__(kBlock) void k_tensor_ab(const StepStatic P, const float* inv, int w, float* out) {
  const LossK& L = P.loss[w];
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (q >= L.cells) return;
  const long long r0 = q / L.bz;
  const int k = (int)(q - r0 * L.bz) + L.lo[2];
  const int j = (int)(r0 % L.by) + L.lo[1];
  const int i = (int)(r0 / L.by) + L.lo[0];
  const long long o = ((long long)i * P.ny + j) * P.nz + k;
  double iv[9], sg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = 0; t < 9; ++t) iv[t] = (double)inv[(long long)t * P.N + o];
  for (int t = 0; t < 3; ++t) sg[4 * t] = (double)L.sig[(L.comps == 3 ? t : 0) * L.cells + q];
  if (L.comps == 9)
    for (int t = 0; t < 9; ++t) sg[t] = (double)L.sig[(long long)t * L.cells + q];
  const double eta = w == 0 ? (double)P.eta0 : 1.0 / (double)P.eta0;
  const double f = (double)P.c * eta / 2.0;
  double is[9], M1[9], M2[9], R1[9], R2[9], X[9];
  mul3(iv, sg, is);
  for (int t = 0; t < 9; ++t) {
    const double e = (t % 4 == 0) ? 1.0 : 0.0;
    M1[t] = e + f * is[t];
    M2[t] = e - f * is[t];
  }
  inverse3(M1, R1);
  inverse3(M2, R2);
  mul3(R1, M2, X);
  for (int t = 0; t < 9; ++t) out[(long long)t * L.cells + q] = (float)X[t];
  mul3(R1, iv, X);
  for (int t = 0; t < 9; ++t) out[(long long)(9 + t) * L.cells + q] = (float)((double)P.c * X[t]);
  mul3(R2, M1, X);
  for (int t = 0; t < 9; ++t) out[(long long)(18 + t) * L.cells + q] = (float)X[t];
  mul3(R2, iv, X);
  for (int t = 0; t < 9; ++t) out[(long long)(27 + t) * L.cells + q] = (float)((double)P.c * X[t]);
}

hipError_t launch_tensor_ab(const StepStatic& P, const StepDyn& D, hipStream_t s) {
  for (int w = 0; w < 2; ++w) {
    if (!P.tsig[w]) continue;
    const unsigned g = (unsigned)((P.loss[w].cells + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_tensor_ab, dim3(g), dim3(kBlock), 0, s, P, w == 0 ? D.ie : D.im, w, (float*)D.ab[w]);
  }
  return hipGetLastError();
}

static hipError_t launch_tensor(const StepStatic& P, const StepDyn& D, int field, int reverse, hipStream_t s) {
  const bool x = extras(P);
  const TensorCurl fc = x ? k_tensor_curl<true> : k_tensor_curl<false>;
  const TensorSweep ft = field == 0 ? (x ? k_tensor_E<true> : k_tensor_E<false>)
                                    : (x ? k_tensor_H<true> : k_tensor_H<false>);
  hipLaunchKernelGGL(fc, grid_of(P), dim3(kBlock), 0, s, P, D, field, reverse ? 0 : 1, reverse ? 0 : 1);
  hipLaunchKernelGGL(ft, grid_of(P), dim3(kBlock), 0, s, P, D, reverse);
  return hipGetLastError();
}

hipError_t launch_update_E(const StepStatic& P, const StepDyn& D, int ec, hipStream_t s) {
  if (ec == 9) return launch_tensor(P, D, 0, 0, s);
  const bool x = extras(P);
  SweepKernel f;
  if (ec == 3) f = x ? k_update_E<3, true> : k_update_E<3, false>;
  else f = x ? k_update_E<1, true> : k_update_E<1, false>;
  hipLaunchKernelGGL(f, grid_of(P), dim3(kBlock), 0, s, P, D);
  return hipGetLastError();
}

hipError_t launch_update_H(const StepStatic& P, const StepDyn& D, int mc, hipStream_t s) {
  if (mc == 9) return launch_tensor(P, D, 1, 0, s);
  const bool x = extras(P);
  SweepKernel f;
  if (mc == 3) f = x ? k_update_H<3, true> : k_update_H<3, false>;
  else if (mc == 1) f = x ? k_update_H<1, true> : k_update_H<1, false>;
  else f = x ? k_update_H<0, true> : k_update_H<0, false>;
  hipLaunchKernelGGL(f, grid_of(P), dim3(kBlock), 0, s, P, D);
  return hipGetLastError();
}

hipError_t launch_update_E4(const StepStatic& P, const StepDyn& D, int ec, hipStream_t s) {
  const bool x = extras(P);
  SweepKernel f;
  SweepKernel fs;
  const bool c = P.cax != 0;
  if (ec == 3) {
    f = x ? (c ? k_update_E4<3, true, true> : k_update_E4<3, true, false>)
          : (c ? k_update_E4<3, false, true> : k_update_E4<3, false, false>);
    fs = x ? k_update_E4s<3, true> : k_update_E4s<3, false>;
  } else {
    f = x ? (c ? k_update_E4<1, true, true> : k_update_E4<1, true, false>)
          : (c ? k_update_E4<1, false, true> : k_update_E4<1, false, false>);
    fs = x ? k_update_E4s<1, true> : k_update_E4s<1, false>;
  }
  hipLaunchKernelGGL(f, dim3((unsigned)P.vnblocks_pad), dim3(kBlock), 0, s, P, D);
  if (P.vnslow) hipLaunchKernelGGL(fs, dim3((unsigned)(4 * P.vnslow)), dim3(kBlock), 0, s, P, D);
  return hipGetLastError();
}

hipError_t launch_update_H4(const StepStatic& P, const StepDyn& D, int mc, hipStream_t s) {
  const bool x = extras(P);
  SweepKernel f;
  SweepKernel fs;
  const bool c = P.cax != 0;
  if (mc == 3) {
    f = x ? (c ? k_update_H4<3, true, true> : k_update_H4<3, true, false>)
          : (c ? k_update_H4<3, false, true> : k_update_H4<3, false, false>);
    fs = x ? k_update_H4s<3, true> : k_update_H4s<3, false>;
  } else if (mc == 1) {
    f = x ? (c ? k_update_H4<1, true, true> : k_update_H4<1, true, false>)
          : (c ? k_update_H4<1, false, true> : k_update_H4<1, false, false>);
    fs = x ? k_update_H4s<1, true> : k_update_H4s<1, false>;
  } else {
    f = x ? (c ? k_update_H4<0, true, true> : k_update_H4<0, true, false>)
          : (c ? k_update_H4<0, false, true> : k_update_H4<0, false, false>);
    fs = x ? k_update_H4s<0, true> : k_update_H4s<0, false>;
  }
  hipLaunchKernelGGL(f, dim3((unsigned)P.vnblocks_pad), dim3(kBlock), 0, s, P, D);
  if (P.vnslowH) hipLaunchKernelGGL(fs, dim3((unsigned)(4 * P.vnslowH)), dim3(kBlock), 0, s, P, D);
  return hipGetLastError();
}

hipError_t launch_record(const StepStatic& P, const StepDyn& D, long long iface_cells, hipStream_t s) {
  if (iface_cells > 0) {
    hipLaunchKernelGGL(k_record, dim3((unsigned)((iface_cells + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, P, D);
  }
  return hipGetLastError();
}

hipError_t launch_restore(const StepStatic& P, const StepDyn& D, long long iface_cells, hipStream_t s) {
  if (iface_cells > 0) {
    hipLaunchKernelGGL(k_restore, dim3((unsigned)((iface_cells + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, P, D);
  }
  return hipGetLastError();
}

// k_reverse_H for diagonal inverse permeability of mc components, per-axis CPML under acax
static void (*reverse_H_kernel(const StepStatic& P, int mc))(const StepStatic, const StepDyn) {
  if (P.acax) return mc == 3 ? k_reverse_H<3, true> : (mc == 1 ? k_reverse_H<1, true> : k_reverse_H<0, true>);
  return mc == 3 ? k_reverse_H<3, false> : (mc == 1 ? k_reverse_H<1, false> : k_reverse_H<0, false>);
}

hipError_t launch_reverse_H(const StepStatic& P, const StepDyn& D, int mc, hipStream_t s) {
  hipLaunchKernelGGL(reverse_H_kernel(P, mc), grid_of(P), dim3(kBlock), 0, s, P, D);
  return hipGetLastError();
}

hipError_t launch_reverse(const StepStatic& P, const StepDyn& D, int ec, int mc, hipStream_t s) {
  if (mc == 9) {
    const hipError_t e = launch_tensor(P, D, 1, 1, s);
    if (e != hipSuccess) return e;
  } else {
    hipLaunchKernelGGL(reverse_H_kernel(P, mc), grid_of(P), dim3(kBlock), 0, s, P, D);
  }
  if (ec == 9) return launch_tensor(P, D, 0, 1, s);
  if (ec == 3) hipLaunchKernelGGL(k_reverse_E<3>, grid_of(P), dim3(kBlock), 0, s, P, D);
  else hipLaunchKernelGGL(k_reverse_E<1>, grid_of(P), dim3(kBlock), 0, s, P, D);
  return hipGetLastError();
}

// ---------------------------------------------------------------- multi-instance batch
// One tile per block over the concatenated tile space of the batch. Block h serves XCD h % 8 and
// takes tile (h % 8) * per + h / 8, so each XCD sweeps a contiguous eighth and neighbouring planes
// of one instance share an L2; blocks past the tile count (the grid is rounded up) exit.
__device__ __forceinline__ bool batch_tile(const BatchTable* __restrict__ B, long long& t, int& e) {
  const int ni = B->ninst;
  const long long total = B->tile_start[ni];
  const long long per = (total + kXcds - 1) / kXcds;
  const long long w = blockIdx.x / kXcds;
  t = (long long)(blockIdx.x % kXcds) * per + w;
  if (w >= per || t >= total) return false;
  int lo = 0, hi = ni - 1;  // the entry e with tile_start[e] <= t < tile_start[e + 1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (B->tile_start[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  e = lo;
  return true;
}

template <int EC>
__global__ __launch_bounds__(kBlock) void k_batch_E(const BatchTable* __restrict__ B, const int* __restrict__ base,
                                                    int r) {
  long long t;
  int e;
  if (!batch_tile(B, t, e)) return;
  const BatchInst& I = B->inst[e];
  const int n = base[I.slot] + r;
  if (n >= I.stop) return;
  update_E_tile<EC, true, true>(I.P, I.D, n, t - B->tile_start[e]);
}

template <int MC>
__global__ __launch_bounds__(kBlock) void k_batch_H(const BatchTable* __restrict__ B, const int* __restrict__ base,
                                                    int r) {
  long long t;
  int e;
  if (!batch_tile(B, t, e)) return;
  const BatchInst& I = B->inst[e];
  const int n = base[I.slot] + r;
  if (n >= I.stop) return;
  const int rs = I.P.rec_slot ? I.P.rec_slot[n] : n;
  char* record = (I.rec_base && rs >= 0) ? I.rec_base + (long long)rs * I.P.rec_slot_bytes : nullptr;
  update_H_tile<MC, true, true>(I.P, I.D, n, record, t - B->tile_start[e]);
}

__global__ void k_batch_set(int* base, int* flags, const BatchSet set) {
  const int t = threadIdx.x;
  if (t < set.count) {
    base[set.slot[t]] = set.value[t];
    flags[set.slot[t]] = 0;
  }
}

__global__ void k_batch_advance(const BatchTable* __restrict__ B, int* base, int steps) {
  const int e = threadIdx.x;
  if (e >= B->ninst) return;
  const BatchInst& I = B->inst[e];
  const int n = base[I.slot];
  base[I.slot] = n + steps < I.stop ? n + steps : (n > I.stop ? n : I.stop);
}

int batch_sweep_grid(long long tiles) {
  // at least 8 * ceil(tiles / 8) blocks, rounded up to a multiple of 2^(floor(log2 tiles) - 3) so
  // that a changing batch reuses few distinct grids (at most 12.5 % idle blocks)
  long long g = ((tiles + kXcds - 1) / kXcds) * kXcds;
  if (g < kXcds) g = kXcds;
  int lg = 0;
  while ((1ll << (lg + 1)) <= g) ++lg;
  const long long q = std::max<long long>(kXcds, 1ll << std::max(0, lg - 3));
  g = ((g + q - 1) / q) * q;
  return (int)g;
}

hipError_t launch_batch_set(int* base, int* flags, const BatchSet& set, hipStream_t s) {
  if (set.count > 0) hipLaunchKernelGGL(k_batch_set, dim3(1), dim3(kMaxBatch), 0, s, base, flags, set);
  return hipGetLastError();
}

hipError_t launch_batch_sweeps(const BatchTable* tab, const int* base, int r, int ec, int mc, int grid,
                               hipStream_t s) {
  if (ec == 3) hipLaunchKernelGGL(k_batch_E<3>, dim3(grid), dim3(kBlock), 0, s, tab, base, r);
  else hipLaunchKernelGGL(k_batch_E<1>, dim3(grid), dim3(kBlock), 0, s, tab, base, r);
  if (mc == 3) hipLaunchKernelGGL(k_batch_H<3>, dim3(grid), dim3(kBlock), 0, s, tab, base, r);
  else if (mc == 1) hipLaunchKernelGGL(k_batch_H<1>, dim3(grid), dim3(kBlock), 0, s, tab, base, r);
  else hipLaunchKernelGGL(k_batch_H<0>, dim3(grid), dim3(kBlock), 0, s, tab, base, r);
  return hipGetLastError();
}

hipError_t launch_batch_advance(const BatchTable* tab, int* base, int steps, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(kMaxBatch), 0, s, tab, base, steps);
  return hipGetLastError();
}

}  // namespace fdtdx
