// Device helpers shared by the two-sweep and the fused single-pass kernels: raw-difference curl
// derivatives, the CPML correction (curl.py:284-308, 371-394; PML.py:138-190) and the source hooks
// (update.py:494-519, 824-849). All arithmetic keeps the reference's float32 expression order.
#pragma once
#include "engine_common.h"

// Descriptor loops (slabs, sources) are rolled or fully unrolled per kernel through the template
// flag U. Measured on the 400x400x200 forward: the single-solve sweeps (descriptors in the kernel
// argument block) run 0.66 ms/step rolled against 1.12 ms/step unrolled, while the batch sweeps
// (descriptors read from the device table) run 0.68 ms/step unrolled against 0.90 rolled.

namespace fdtdx {

// Derivatives feeding the two psi of a slab on `axis` (curl.py:289-294 for E, 375-380 for H):
//   curl H: x -> (dxHz, dxHy), y -> (dyHx, dyHz), z -> (dzHy, dzHx)
//   curl E: x -> (dxEz, dxEy), y -> (dyEx, dyEz), z -> (dzEy, dzEx)
// and the scatter curl[(a+1)%3] += -corr1, curl[(a+2)%3] += corr2.
struct Deriv {
  float dx_x, dx_y, dx_z;  // d/dx of the x, y, z components
  float dy_x, dy_y, dy_z;
  float dz_x, dz_y, dz_z;
};

__device__ __forceinline__ void pml_pair(int axis, const Deriv& d, float& d1, float& d2) {
  if (axis == 0) { d1 = d.dx_z; d2 = d.dx_y; }
  else if (axis == 1) { d1 = d.dy_x; d2 = d.dy_z; }
  else { d1 = d.dz_y; d2 = d.dz_x; }
}

// Written as selects (adding an exact 0.0f to the untouched component) so the compiler keeps the
// three curl components in registers instead of an axis-indexed stack array.
__device__ __forceinline__ void scatter(int axis, float c1, float c2, float& cx, float& cy, float& cz) {
  const float ax_ = axis == 1 ? c2 : (axis == 2 ? -c1 : 0.f);
  const float ay_ = axis == 0 ? -c1 : (axis == 2 ? c2 : 0.f);
  const float az_ = axis == 0 ? c2 : (axis == 1 ? -c1 : 0.f);
  cx = cx + ax_;
  cy = cy + ay_;
  cz = cz + az_;
}

__device__ __forceinline__ long long slab_index(const SlabK& s, int i, int j, int k) {
  int li = i, lj = j, lk = k;
  if (s.axis == 0) li -= s.lo; else if (s.axis == 1) lj -= s.lo; else lk -= s.lo;
  return ((long long)li * s.sy + lj) * s.sz + lk;
}

// Apply the CPML of every slab containing (i,j,k). which = 0 -> E coefficients (curl H),
// 1 -> H coefficients (curl E). update = 1: psi is advanced from psi_in and stored to psi_out;
// update = 2: advanced but not stored (a recomputed halo cell); update = 0: psi read frozen
// (simulate_boundaries=False, PML.py:181-182).
__device__ __forceinline__ void cpml_slab(const SlabK& sl, const float* psi_in, float* psi_out, int which,
                                          int update, int i, int j, int k, const Deriv& d, float& cx, float& cy,
                                          float& cz) {
  const int coord = sl.axis == 0 ? i : (sl.axis == 1 ? j : k);
  if (coord < sl.lo || coord >= sl.hi) return;
  const int dep = coord - sl.lo;
  const float* cf = sl.coef + (which ? 3 * sl.L : 0);
  const float b = cf[dep], a = cf[sl.L + dep];
  float d1, d2;
  pml_pair(sl.axis, d, d1, d2);
  const long long q = slab_index(sl, i, j, k);
  const long long off = sl.psi_off + (which ? 2 * sl.cells : 0) + q;
  float p1 = ldg(psi_in + off), p2 = ldg(psi_in + off + sl.cells);
  if (update) {
    p1 = b * p1 + a * d1;
    p2 = b * p2 + a * d2;
    if (update == 1) {
      stg(psi_out + off, p1);
      stg(psi_out + off + sl.cells, p2);
    }
  }
  float c1 = p1, c2 = p2;
  if (!sl.kappa_one) {
    const float ikm1 = cf[2 * sl.L + dep];
    c1 = ikm1 * d1 + p1;
    c2 = ikm1 * d2 + p2;
  }
  scatter(sl.axis, c1, c2, cx, cy, cz);
}

// ---------------------------------------------------------------- per-axis CPML (acax)
// Along each axis a cell lies in at most one slab (StepStatic::acax); the side holding it follows
// from the coordinate and the two sides' descriptors (wave-uniform, selected per lane), so the psi
// and coefficient loads of every axis can be issued before any value is used, instead of behind
// the slab loop's branches.

// Float offset of the first psi component of the cell at depth dep of a slab on axis AX (whose
// psi block starts at base; W: the H-type pair, after the E-type pair) at transverse position
// (i, j, k); the second component follows at + cells.
template <int AX>
__device__ __forceinline__ long long ax_psi(const StepStatic& P, long long base, int W, long long cells, int L,
                                            int dep, int i, int j, int k) {
  const long long b = base + (W ? 2 * cells : 0);
  if (AX == 0) return b + (long long)dep * P.plane + (long long)j * P.nz + k;
  if (AX == 1) return b + ((long long)i * L + dep) * P.nz + k;
  return b + ((long long)i * P.ny + j) * L + dep;
}

__device__ __forceinline__ int ax_side_of(const StepStatic::AxSlab& A, const StepStatic::AxSlab& B, int c) {
  return (c >= A.lo && c < A.hi) ? 0 : ((c >= B.lo && c < B.hi) ? 1 : -1);
}

// One axis at one cell: psi of the cell's slab (W-type pair) and that slab's coefficients (b and a
// only with BA).
struct AxCell {
  int s;                // side of the slab holding the cell (-1: none)
  bool kap;             // kappa != 1
  long long q, cells;   // psi offset (first component) and the slab's cell count
  float p1, p2;         // psi of the two components
  float b, a, km1;      // coefficients at the cell's depth (km1 = 1/kappa - 1, with kap only)
};

template <int W, int AX, bool BA = true>
__device__ __forceinline__ void ax_cell(const StepStatic& P, const float* psi, int i, int j, int k, AxCell& t) {
  const StepStatic::AxSlab& A = P.axd[AX][0];
  const StepStatic::AxSlab& B = P.axd[AX][1];
  const int c = AX == 0 ? i : (AX == 1 ? j : k);
  t.s = ax_side_of(A, B, c);
  t.kap = false;
  t.q = 0;
  t.cells = 0;
  t.p1 = t.p2 = t.b = t.a = t.km1 = 0.f;
  if (t.s < 0) return;
  const bool h = t.s == 1;
  const int L = h ? B.L : A.L, dep = c - (h ? B.lo : A.lo);
  t.cells = h ? B.cells : A.cells;
  t.kap = !(h ? B.kappa_one : A.kappa_one);
  t.q = ax_psi<AX>(P, h ? B.psi_off : A.psi_off, W, t.cells, L, dep, i, j, k);
  const float* cf = (h ? B.coef : A.coef) + (W ? 3 * L : 0);
  t.p1 = psi[t.q];
  t.p2 = psi[t.q + t.cells];
  if (BA) {
    t.b = cf[dep];
    t.a = cf[L + dep];
  }
  if (t.kap) t.km1 = cf[2 * L + dep];
}

// The frozen-psi correction (cpml_slab with update = 0) of the preloaded slab cells of the three
// axes, added x, y, z.
__device__ __forceinline__ void cpml_axes_frozen(const AxCell* ax, const Deriv& d, float& cx, float& cy, float& cz) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (ax[a].s < 0) continue;
    float d1, d2;
    pml_pair(a, d, d1, d2);
    float c1 = ax[a].p1, c2 = ax[a].p2;
    if (ax[a].kap) {
      c1 = ax[a].km1 * d1 + ax[a].p1;
      c2 = ax[a].km1 * d2 + ax[a].p2;
    }
    scatter(a, c1, c2, cx, cy, cz);
  }
}

// `mask`: tile mask (bit s = slab s may touch the tile); slabs in object order.
template <bool U>
__device__ __forceinline__ void apply_cpml(const StepStatic& P, uint32_t mask, const float* psi_in, float* psi_out,
                                           int which, int update, int i, int j, int k, const Deriv& d, float& cx,
                                           float& cy, float& cz) {
  if (!(mask & ((1u << FDTDX_MAX_CPML) - 1u))) return;
  if constexpr (U) {
#pragma unroll
    for (int s = 0; s < FDTDX_MAX_CPML; ++s) {
      if (s >= P.nslab) break;
      if (!((mask >> s) & 1u)) continue;
      cpml_slab(P.slab[s], psi_in, psi_out, which, update, i, j, k, d, cx, cy, cz);
    }
  } else {
    uint32_t m = mask & ((1u << P.nslab) - 1u);
    while (m) {
      const int s = __builtin_ctz(m);
      m &= m - 1u;
      cpml_slab(P.slab[s], psi_in, psi_out, which, update, i, j, k, d, cx, cy, cz);
    }
  }
}

// Source hooks of one field at step n; inverse flips the sign (update_E_reverse/update_H_reverse).
__device__ __forceinline__ void source_one(const StepStatic& P, const SrcK& sr, const float* ie, const float* im,
                                           int field, int n, bool inverse, int i, int j, int k, long long o, int ec,
                                           int mc, float& vx, float& vy, float& vz) {
  if (sr.field != field) return;
  if (!in_box(i, j, k, sr.lo, sr.hi)) return;
  if (sr.kind == FDTDX_SRC_POINT) {
    const float v = sr.table[n];
    const float sv = inverse ? -v : v;  // x - v == x + (-v) exactly
    vx = vx + (sr.comp == 0 ? sv : 0.f);
    vy = vy + (sr.comp == 1 ? sv : 0.f);
    vz = vz + (sr.comp == 2 ? sv : 0.f);
  } else {
    const long long pidx = ((long long)(i - sr.lo[0]) * sr.py + (j - sr.lo[1])) * sr.pz + (k - sr.lo[2]);
    for (int t2 = 0; t2 < 2; ++t2) {
      const int tc = sr.target[t2];
      const int sc = sr.src[t2];
      float mult;
      if (field == 0) mult = ie[(ec == 3 ? tc : 0) * P.N + o];
      else mult = mc ? im[(mc == 3 ? tc : 0) * P.N + o] : 1.0f;
      const float term = ((sr.profile[sc * sr.pcells + pidx] * sr.table[sc * P.T + n]) * P.c) * mult;
      const float inc = (inverse ? -sr.sign[t2] : sr.sign[t2]) * term;
      vx = vx + (tc == 0 ? inc : 0.f);
      vy = vy + (tc == 1 ? inc : 0.f);
      vz = vz + (tc == 2 ? inc : 0.f);
    }
  }
}

// `mask`: tile mask (bit kMaskSrc + s = source s may touch the tile); sources in list order.
template <bool U>
__device__ __forceinline__ void apply_sources(const StepStatic& P, uint32_t mask, const float* ie, const float* im,
                                              int field, int n, bool inverse, int i, int j, int k,
                                              long long o, int ec, int mc, float& vx, float& vy, float& vz) {
  if constexpr (U) {
#pragma unroll
    for (int s = 0; s < FDTDX_MAX_SOURCES; ++s) {
      if (s >= P.nsrc) break;
      if (!((mask >> (kMaskSrc + s)) & 1u)) continue;
      source_one(P, P.src[s], ie, im, field, n, inverse, i, j, k, o, ec, mc, vx, vy, vz);
    }
  } else {
    uint32_t m = (mask >> kMaskSrc) & ((1u << P.nsrc) - 1u);
    while (m) {
      const int s = __builtin_ctz(m);
      m &= m - 1u;
      source_one(P, P.src[s], ie, im, field, n, inverse, i, j, k, o, ec, mc, vx, vy, vz);
    }
  }
}


// PEC / PMC faces of kind `kind` (update.py:48-89): zero the components tangential to the face on
// its layer. `mask` is the tile mask.
__device__ __forceinline__ void apply_walls(const StepStatic& P, uint32_t mask, int kind, int i, int j, int k,
                                            float& vx, float& vy, float& vz) {
  uint32_t m = (mask >> kMaskWall) & ((1u << P.nwall) - 1u);
  while (m) {
    const int w = __builtin_ctz(m);
    m &= m - 1u;
    const WallK& W = P.wall[w];
    if (W.kind != kind) continue;
    const int coord = W.axis == 0 ? i : (W.axis == 1 ? j : k);
    if (coord != W.index) continue;
    if (W.axis != 0) vx = 0.f;
    if (W.axis != 1) vy = 0.f;
    if (W.axis != 2) vz = 0.f;
  }
}

// True when component c of the field at (i, j, k) is held at zero by a face of kind `kind`.
__device__ __forceinline__ bool wall_zero(const StepStatic& P, uint32_t mask, int kind, int c, int i, int j, int k) {
  uint32_t m = (mask >> kMaskWall) & ((1u << P.nwall) - 1u);
  while (m) {
    const int w = __builtin_ctz(m);
    m &= m - 1u;
    const WallK& W = P.wall[w];
    if (W.kind != kind || W.axis == c) continue;
    const int coord = W.axis == 0 ? i : (W.axis == 1 ? j : k);
    if (coord == W.index) return true;
  }
  return false;
}

// Loss x of the three components at (i, j, k) for conductivity box w (0: electric, 1: magnetic):
// x = c*sigma*eta0*inv/2 (update.py:300-305) or x = c*sigma/eta0*inv/2 (update.py:743-746), in
// that float order, with inv the cell's inverse permittivity or permeability. False (and x
// untouched) where nothing conducts.
__device__ __forceinline__ bool loss_x(const StepStatic& P, uint32_t mask, int w, int i, int j, int k,
                                       float inv0, float inv1, float inv2, float* x) {
  const LossK& L = P.loss[w];
  if (!((mask >> (kMaskLoss + w)) & 1u) || !in_box(i, j, k, L.lo, L.hi)) return false;
  const long long q = ((long long)(i - L.lo[0]) * L.by + (j - L.lo[1])) * L.bz + (k - L.lo[2]);
  const float s0 = L.sig[q];
  const float s1 = L.comps == 3 ? L.sig[L.cells + q] : s0;
  const float s2 = L.comps == 3 ? L.sig[2 * L.cells + q] : s0;
  const float c = P.c;
  if (w == 0) {
    x[0] = (((c * s0) * P.eta0) * inv0) / 2.0f;
    x[1] = (((c * s1) * P.eta0) * inv1) / 2.0f;
    x[2] = (((c * s2) * P.eta0) * inv2) / 2.0f;
  } else {
    x[0] = (((c * s0) / P.eta0) * inv0) / 2.0f;
    x[1] = (((c * s1) / P.eta0) * inv1) / 2.0f;
    x[2] = (((c * s2) / P.eta0) * inv2) / 2.0f;
  }
  return true;
}

// Electric conductivity c*sigma*eta0 of the three components (twice the loss per unit inverse
// permittivity), for the adjoint's inv_eps derivative. False where nothing conducts.
__device__ __forceinline__ bool loss_a2(const StepStatic& P, uint32_t mask, int i, int j, int k, float* a2) {
  const LossK& L = P.loss[0];
  if (!((mask >> kMaskLoss) & 1u) || !in_box(i, j, k, L.lo, L.hi)) return false;
  const long long q = ((long long)(i - L.lo[0]) * L.by + (j - L.lo[1])) * L.bz + (k - L.lo[2]);
  for (int cc = 0; cc < 3; ++cc) a2[cc] = (P.c * L.sig[(L.comps == 3 ? cc : 0) * L.cells + q]) * P.eta0;
  return true;
}

}  // namespace fdtdx
