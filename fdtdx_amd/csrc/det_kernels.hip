// Detector kernels: forward accumulation (energy, field snapshot, Poynting flux, on-the-fly DFT)
// and the transposed detector update that injects loss cotangents into the adjoint fields.
//
// Detectors see E co-located at the Ez point and the time-centred H = (H_prev + H)/2 interpolated
// to the same point (update.py:1084-1102, core/physics/curl.py:86-224); fields beyond the domain
// are zero (pad_fields, core/misc.py:615-641).
#include "engine_common.h"
#include "kernels.h"

namespace fdtdx {

struct Src3 {  // a 3-component field sampled either from a full array or from a box snapshot
  const float* p;
  long long stride;  // component stride
  int lo[3], hi[3];  // box (full domain for full arrays)
  int sy, sz;
  int h;             // 0: E, 1: H (the mirror parity below an electric symmetry plane)
  // complex-field plans: the other part of the field in the same layout, and which part p holds
  // (+1 real, -1 imaginary; 0 for real plans)
  const float* q;
  int cp;
};

// Below the min face of an electric symmetry plane on axis a (pad_fields_with_symmetry_mirror,
// update.py:139-190, core/physics/symmetry.py:18-108): components sampled on the plane
// (tangential E, normal H) are odd and mirror row 1, the others are even and mirror row 0. Axes are
// folded in order, so a corner takes both signs.
__device__ __forceinline__ float mirror_fold(const StepStatic& P, int h, int c, int& i, int& j, int& k) {
  float sign = 1.f;
  int* q[3] = {&i, &j, &k};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!P.mirror[a] || *q[a] != -1) continue;
    const bool on = h ? c == a : c != a;
    *q[a] = on ? 1 : 0;
    if (on) sign = -sign;
  }
  return sign;
}

// Branch-free: a point outside the domain reads the nearest cell of the source box and is
// multiplied by zero, so every load of an interpolation is issued unconditionally and all of them
// are in flight together.
// A point beyond a periodic face reads the wrapped cell; in a complex-field plan with Bloch phases
// it reads that cell's complex value times exp(-i theta_min) (beyond the min face) or
// exp(+i theta_max) (beyond the max face), per wrapped axis (pad_fields_for_boundaries,
// update.py:92-137, apply_pad_correction, bloch.py:61-97).
__device__ __forceinline__ float at(const StepStatic& P, const Src3& s, int c, int i, int j, int k) {
  const float sign = (P.mirror[0] | P.mirror[1] | P.mirror[2]) ? mirror_fold(P, s.h, c, i, j, k) : 1.f;
  float C = 1.f, S = 0.f;  // the accumulated phase exp(i t) = C + i S of the wrapped axes
  if (s.cp && P.bloch) {
    const int q3[3] = {i, j, k};
    const int n3[3] = {P.nx, P.ny, P.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!P.wrap[a] || (q3[a] >= 0 && q3[a] < n3[a])) continue;
      const int side = q3[a] < 0 ? 0 : 1;
      const float ca = P.bcos[a][side], sa = side == 0 ? -P.bsin[a][side] : P.bsin[a][side];
      const float C2 = C * ca - S * sa;
      S = C * sa + S * ca;
      C = C2;
    }
  }
  if (P.wrap[0]) i = wrap_coord(i, P.nx);
  if (P.wrap[1]) j = wrap_coord(j, P.ny);
  if (P.wrap[2]) k = wrap_coord(k, P.nz);
  const bool in = i >= 0 && j >= 0 && k >= 0 && i < P.nx && j < P.ny && k < P.nz;
  const int ci = min(max(i, s.lo[0]), s.hi[0] - 1);
  const int cj = min(max(j, s.lo[1]), s.hi[1] - 1);
  const int ck = min(max(k, s.lo[2]), s.hi[2] - 1);
  const long long idx = c * s.stride + ((long long)(ci - s.lo[0]) * s.sy + (cj - s.lo[1])) * s.sz + (ck - s.lo[2]);
  float v = s.p[idx];
  if (S != 0.f) v = C * v - (float)s.cp * S * s.q[idx];  // this part of (C + i S) z
  else if (C != 1.f) v = C * v;
  return in ? sign * v : 0.f;
}

__device__ __forceinline__ Src3 full3(const StepStatic& P, const float* p, int h) {
  Src3 s;
  s.q = nullptr;
  s.cp = 0;
  s.h = h;
  s.p = p; s.stride = P.N;
  s.lo[0] = s.lo[1] = s.lo[2] = 0;
  s.hi[0] = P.nx; s.hi[1] = P.ny; s.hi[2] = P.nz;
  s.sy = P.ny; s.sz = P.nz;
  return s;
}

__device__ __forceinline__ Src3 box3(const float* p, const int* lo, const int* hi, int h) {
  Src3 s;
  s.q = nullptr;
  s.cp = 0;
  s.h = h;
  s.p = p;
  s.sy = hi[1] - lo[1]; s.sz = hi[2] - lo[2];
  s.stride = (long long)(hi[0] - lo[0]) * s.sy * s.sz;
  for (int a = 0; a < 3; ++a) { s.lo[a] = lo[a]; s.hi[a] = hi[a]; }
  return s;
}

// Time-centred H at a point: (H_prev + H) / 2 (update.py:1088, 1098).
__device__ __forceinline__ float hc(const StepStatic& P, const Src3& hp, const Src3& h, int c, int i, int j, int k) {
  return (at(P, hp, c, i, j, k) + at(P, h, c, i, j, k)) / 2.0f;
}

// The stencil of interp at (i, j, k) reads i-1, j-1 and k+1: when that stays inside a source's box
// (which lies inside the domain) no halo, wrap, mirror or Bloch phase applies and the reads are
// plain offsets.
__device__ __forceinline__ bool stencil_inside(const Src3& s, int i, int j, int k) {
  return i - 1 >= s.lo[0] && j - 1 >= s.lo[1] && k + 1 < s.hi[2] && i < s.hi[0] && j < s.hi[1] && k >= s.lo[2];
}

__device__ __forceinline__ long long src_index(const Src3& s, int i, int j, int k) {
  return ((long long)(i - s.lo[0]) * s.sy + (j - s.lo[1])) * s.sz + (k - s.lo[2]);
}

// interp where stencil_inside holds for e, hp and h: the same expressions in the same order, so
// the result is identical to the general path's.
__device__ __forceinline__ void interp_inside(const Src3& e, const Src3& hp, const Src3& h, int i, int j, int k,
                                              float* Ei, float* Hi) {
  const long long oe = src_index(e, i, j, k);
  const long long xe = (long long)e.sy * e.sz, ye = e.sz;
  const float* ex = e.p;
  const float* ey = e.p + e.stride;
  const float* ez = e.p + 2 * e.stride;
  const float lo_x = 0.5f * (ex[oe] + ex[oe - xe]);
  const float up_x = 0.5f * (ex[oe + 1] + ex[oe - xe + 1]);
  Ei[0] = (lo_x + up_x) / 2.0f;
  const float lo_y = 0.5f * (ey[oe] + ey[oe - ye]);
  const float up_y = 0.5f * (ey[oe + 1] + ey[oe - ye + 1]);
  Ei[1] = (lo_y + up_y) / 2.0f;
  Ei[2] = ez[oe];
  const long long op = src_index(hp, i, j, k), oh = src_index(h, i, j, k);
  const long long xp = (long long)hp.sy * hp.sz, yp = hp.sz, xh = (long long)h.sy * h.sz, yh = h.sz;
  // time-centred H at the offset (dx, dy, dz) in {-1, 0} x {-1, 0} x {0, 1}
  auto hcq = [&](int c, int dx, int dy, int dz) {
    const float a = hp.p[c * hp.stride + op + dx * xp + dy * yp + dz];
    const float b = h.p[c * h.stride + oh + dx * xh + dy * yh + dz];
    return (a + b) / 2.0f;
  };
  Hi[0] = 0.5f * (hcq(0, 0, 0, 0) + hcq(0, 0, -1, 0));
  Hi[1] = 0.5f * (hcq(1, 0, 0, 0) + hcq(1, -1, 0, 0));
  float zz[2];
#pragma unroll
  for (int dz = 0; dz < 2; ++dz) {
    const float a = 0.5f * (hcq(2, 0, 0, dz) + hcq(2, -1, 0, dz));
    const float b = 0.5f * (hcq(2, 0, -1, dz) + hcq(2, -1, -1, dz));
    zz[dz] = 0.5f * (a + b);
  }
  Hi[2] = (zz[0] + zz[1]) / 2.0f;
}

// interpolate_fields (curl.py:86-224) at cell (i,j,k) on a uniform grid.
__device__ __forceinline__ void interp(const StepStatic& P, const Src3& e, const Src3& hp, const Src3& h, int i, int j, int k,
                       float* Ei, float* Hi) {
  if (stencil_inside(e, i, j, k) && stencil_inside(hp, i, j, k) && stencil_inside(h, i, j, k)) {
    interp_inside(e, hp, h, i, j, k, Ei, Hi);
    return;
  }
  const float lo_x = 0.5f * (at(P, e, 0, i, j, k) + at(P, e, 0, i - 1, j, k));
  const float up_x = 0.5f * (at(P, e, 0, i, j, k + 1) + at(P, e, 0, i - 1, j, k + 1));
  Ei[0] = (lo_x + up_x) / 2.0f;
  const float lo_y = 0.5f * (at(P, e, 1, i, j, k) + at(P, e, 1, i, j - 1, k));
  const float up_y = 0.5f * (at(P, e, 1, i, j, k + 1) + at(P, e, 1, i, j - 1, k + 1));
  Ei[1] = (lo_y + up_y) / 2.0f;
  Ei[2] = at(P, e, 2, i, j, k);
  Hi[0] = 0.5f * (hc(P, hp, h, 0, i, j, k) + hc(P, hp, h, 0, i, j - 1, k));
  Hi[1] = 0.5f * (hc(P, hp, h, 1, i, j, k) + hc(P, hp, h, 1, i - 1, j, k));
  float zz[2];
  for (int dz = 0; dz < 2; ++dz) {
    const float a = 0.5f * (hc(P, hp, h, 2, i, j, k + dz) + hc(P, hp, h, 2, i - 1, j, k + dz));
    const float b = 0.5f * (hc(P, hp, h, 2, i, j - 1, k + dz) + hc(P, hp, h, 2, i - 1, j - 1, k + dz));
    zz[dz] = 0.5f * (a + b);
  }
  Hi[2] = (zz[0] + zz[1]) / 2.0f;
}

__device__ __forceinline__ float pick6(int c, const float* E, const float* H) {
  return c == 0 ? E[0] : c == 1 ? E[1] : c == 2 ? E[2] : c == 3 ? H[0] : c == 4 ? H[1] : H[2];
}

__device__ __forceinline__ void add6(int c, float v, float* E, float* H) {
  E[0] += c == 0 ? v : 0.f;
  E[1] += c == 1 ? v : 0.f;
  E[2] += c == 2 ? v : 0.f;
  H[0] += c == 3 ? v : 0.f;
  H[1] += c == 4 ? v : 0.f;
  H[2] += c == 5 ? v : 0.f;
}

__device__ __forceinline__ void region_cell(const DetK& K, long long q, int& i, int& j, int& k) {
  const int sy = K.hi[1] - K.lo[1], sz = K.hi[2] - K.lo[2];
  const long long r = q / sz;
  k = (int)(q - r * sz) + K.lo[2];
  const long long ii = r / sy;
  j = (int)(r - ii * sy) + K.lo[1];
  i = (int)ii + K.lo[0];
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();  // sh may be reused by the next call of this block
  return r;
}

// compute_energy (core/physics/metrics.py:56-66) of the co-located fields at cell o.
// |E|^2 and |H|^2 with the imaginary parts Ej, Hj (all zero for real fields, where adding their
// zero squares leaves the real squares unchanged).
__device__ __forceinline__ float energy_of(const StepStatic& P, const DetDyn& D, long long o, const float* Ei,
                                           const float* Hi, const float* Ej, const float* Hj) {
  float eE[3], eH[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float iec = D.ie[(D.ec == 3 ? c : 0) * P.N + o];
    const float e2 = Ei[c] * Ei[c] + Ej[c] * Ej[c];
    const float h2 = Hi[c] * Hi[c] + Hj[c] * Hj[c];
    eE[c] = (0.5f * (1.0f / iec)) * e2;
    const float hfac = D.mc ? 0.5f * (1.0f / D.im[(D.mc == 3 ? c : 0) * P.N + o]) : 0.5f;
    eH[c] = hfac * h2;
  }
  return ((eE[0] + eE[1]) + eE[2]) + ((eH[0] + eH[1]) + eH[2]);
}

// The two views of a complex-field plan's fields for interp: re = true the real part (the other
// part is the imaginary one), else the imaginary part.
__device__ __forceinline__ void complex_views(const StepStatic& P, const DetDyn& D, bool re, Src3& e, Src3& hp,
                                              Src3& h) {
  const float* E = re ? D.E : D.Ei;
  const float* H = re ? D.H : D.Hi;
  const float* Hp = re ? D.Hprev : D.Hprevi;
  e = full3(P, E, 0);
  h = full3(P, H, 1);
  hp = box3(Hp, D.slo, D.shi, 1);
  e.q = re ? D.Ei : D.E;
  h.q = re ? D.Hi : D.H;
  hp.q = re ? D.Hprevi : D.Hprev;
  e.cp = h.cp = hp.cp = re ? 1 : -1;
}

// The two box axes other than `ax`, in increasing order, and the plane index of a cell.
__device__ __forceinline__ void plane_axes(int ax, int& b0, int& b1) {
  b0 = ax == 0 ? 1 : 0;
  b1 = ax == 2 ? 1 : 2;
}

// EnergyDetector(as_slices) with a collapsed-axis mean (energy.py:107-114): thread per cell of the
// plane normal to K.axis, the energy averaged along that axis (float64 sum, then divided), stored
// as slot `slot` of the (n_rec, d0, d1) state.
template <bool CX>
__device__ __forceinline__ void energy_mean_tile(const StepStatic& P, const DetK& K, const DetDyn& D, int slot,
                                                 int blk, int nblk, const Src3& e, const Src3& hp, const Src3& h,
                                                 const Src3& e2, const Src3& hp2, const Src3& h2) {
  constexpr bool cx = CX;
  int b0, b1;
  plane_axes(K.axis, b0, b1);
  const int na = K.hi[K.axis] - K.lo[K.axis];
  const int d1 = K.hi[b1] - K.lo[b1];
  const long long pc = K.cells / na;
  for (long long q = (long long)blk * kBlock + threadIdx.x; q < pc; q += (long long)nblk * kBlock) {
    const int u = (int)(q / d1), v = (int)(q - (long long)u * d1);
    int c[3];
    c[b0] = K.lo[b0] + u;
    c[b1] = K.lo[b1] + v;
    double sum = 0.0;
    for (int t = 0; t < na; ++t) {
      c[K.axis] = K.lo[K.axis] + t;
      const long long o = ((long long)c[0] * P.ny + c[1]) * P.nz + c[2];
      float Ei[3], Hi[3], Ej[3] = {0.f, 0.f, 0.f}, Hj[3] = {0.f, 0.f, 0.f};
      if (K.exact) {
        interp(P, e, hp, h, c[0], c[1], c[2], Ei, Hi);
        if (cx) interp(P, e2, hp2, h2, c[0], c[1], c[2], Ej, Hj);
      } else {
        for (int cc = 0; cc < 3; ++cc) { Ei[cc] = D.E[cc * P.N + o]; Hi[cc] = D.H[cc * P.N + o]; }
        if (cx)
          for (int cc = 0; cc < 3; ++cc) { Ej[cc] = D.Ei[cc * P.N + o]; Hj[cc] = D.Hi[cc * P.N + o]; }
      }
      sum += (double)energy_of(P, D, o, Ei, Hi, Ej, Hj);
    }
    ((float*)D.state)[(long long)slot * pc + q] = (float)(sum / (double)na);
  }
}

// One detector update at step n into record slot `slot`, computed by block blk of nblk blocks
// that stride over the detector's cells. A reducing detector leaves one partial sum per block and
// value: the weighted energy or flux, or the sum of each recorded component (field and phasor
// detectors with reduce_volume, whose box mean is formed in reduce_finalize_tile).
template <bool CX>
__device__ __forceinline__ void detector_tile(const StepStatic& P, const DetK& K, const DetDyn& D, int n, int slot,
                                              int blk, int nblk) {
  // CX: a complex-field plan (D.Ei set): e, hp, h view the real part and e2, hp2, h2 the imaginary one
  constexpr bool cx = CX;
  Src3 e, h, hp, e2, h2, hp2;
  if (cx) {
    complex_views(P, D, true, e, hp, h);
    complex_views(P, D, false, e2, hp2, h2);
  } else {
    e = full3(P, D.E, 0);
    h = full3(P, D.H, 1);
    hp = box3(D.Hprev, D.slo, D.shi, 1);
    e2 = e; h2 = h; hp2 = hp;
  }
  if (K.kind == FDTDX_DET_ENERGY && K.reduce == 2) {
    energy_mean_tile<CX>(P, K, D, slot, blk, nblk, e, hp, h, e2, hp2, h2);
    return;
  }
  const int nv = reduced_values(K);
  double acc[kDetMaxValues] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (long long q = (long long)blk * kBlock + threadIdx.x; q < K.cells; q += (long long)nblk * kBlock) {
    int i, j, k;
    region_cell(K, q, i, j, k);
    const long long o = ((long long)i * P.ny + j) * P.nz + k;
    float Ei[3], Hi[3], Ej[3] = {0.f, 0.f, 0.f}, Hj[3] = {0.f, 0.f, 0.f};
    if (K.exact) {
      interp(P, e, hp, h, i, j, k, Ei, Hi);
      if (cx) interp(P, e2, hp2, h2, i, j, k, Ej, Hj);
    } else {
      for (int c = 0; c < 3; ++c) { Ei[c] = D.E[c * P.N + o]; Hi[c] = D.H[c * P.N + o]; }
      if (cx)
        for (int c = 0; c < 3; ++c) { Ej[c] = D.Ei[c * P.N + o]; Hj[c] = D.Hi[c * P.N + o]; }
    }
    if (K.kind == FDTDX_DET_ENERGY) {  // compute_energy (core/physics/metrics.py:56-66)
      const float en = energy_of(P, D, o, Ei, Hi, Ej, Hj);
      if (nv) acc[0] += (double)(en * K.weight);
      else ((float*)D.state)[(long long)slot * K.cells + q] = en;
    } else if (K.kind == FDTDX_DET_FIELD) {  // field.py:44-62
      for (int c = 0; c < K.ncomp; ++c) {
        const int cc = K.comps[c];
        const float v = pick6(cc, Ei, Hi);
        if (nv) acc[c] += (double)v;
        else ((float*)D.state)[((long long)slot * K.ncomp + c) * K.cells + q] = v;
      }
    } else if (K.kind == FDTDX_DET_FLUX) {  // compute_poynting_flux (metrics.py:99-117)
      // the propagation axis only, or all three components (keep_all_components)
      const int na = K.ncomp == 3 ? 3 : 1;
      for (int t = 0; t < na; ++t) {
        const int a = na == 3 ? t : K.axis;
        const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
        // Re(E x conj H): with complex fields each product adds the imaginary parts' product
        float t1 = pick6(b1, Ei, Hi) * pick6(b2 + 3, Ei, Hi);
        float t2 = pick6(b2, Ei, Hi) * pick6(b1 + 3, Ei, Hi);
        if (cx) {
          t1 = t1 + pick6(b1, Ej, Hj) * pick6(b2 + 3, Ej, Hj);
          t2 = t2 + pick6(b2, Ej, Hj) * pick6(b1 + 3, Ej, Hj);
        }
        float pf = t1 - t2;
        if (K.sign < 0.f) pf = -pf;
        if (nv) acc[t] += (double)(pf * K.weight);
        else ((float*)D.state)[((long long)slot * na + t) * K.cells + q] = pf;
      }
    } else if (nv) {  // reduced PHASOR: the phase factor is uniform over the box, so sum EH first
      for (int c = 0; c < K.ncomp; ++c) acc[c] += (double)pick6(K.comps[c], Ei, Hi);
      if (cx)  // the imaginary parts' sums follow the real parts'
        for (int c = 0; c < K.ncomp; ++c) acc[K.ncomp + c] += (double)pick6(K.comps[c], Ej, Hj);
    } else {  // PHASOR (phasor.py:190-235): state += EH * exp(i w t) * scale * window
      // the state words of up to three frequencies are all loaded before any is updated, so their
      // reads are in flight together (a plane detector launches too few waves to hide them
      // one by one)
      float2* st = (float2*)D.state;
      const float w = K.window[n];
      float v[6], vi[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        v[c] = c < K.ncomp ? pick6(K.comps[c], Ei, Hi) : 0.f;
        vi[c] = c < K.ncomp ? pick6(K.comps[c], Ej, Hj) : 0.f;
      }
      for (int f0 = 0; f0 < K.nfreq; f0 += 3) {
        float2 old[3][6];
#pragma unroll
        for (int df = 0; df < 3; ++df)
#pragma unroll
          for (int c = 0; c < 6; ++c)
            if (f0 + df < K.nfreq && c < K.ncomp) old[df][c] = st[((long long)(f0 + df) * K.ncomp + c) * K.cells + q];
#pragma unroll
        for (int df = 0; df < 3; ++df) {
          const int f = f0 + df;
          if (f >= K.nfreq) break;
          const float cs = K.cos_t[(long long)f * P.T + n];
          const float sn = K.sin_t[(long long)f * P.T + n];
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            if (c >= K.ncomp) break;
            float re = ((v[c] * cs) * K.scale) * w;
            float im = ((v[c] * sn) * K.scale) * w;
            if (cx) {  // (v + i vi)(cos + i sin)
              re = re - ((vi[c] * sn) * K.scale) * w;
              im = ((vi[c] * cs) * K.scale) * w + im;
            }
            float2 s = old[df][c];
            if (K.sign < 0.f) { s.x = s.x - re; s.y = s.y - im; }
            else { s.x = s.x + re; s.y = s.y + im; }
            st[((long long)f * K.ncomp + c) * K.cells + q] = s;
          }
        }
      }
    }
  }
  for (int v = 0; v < kDetMaxValues; ++v) {  // nv is uniform over the block
    if (v >= nv) break;
    const double s = block_sum(acc[v]);
    if (threadIdx.x == 0) D.partial[v * kDetMaxBlocks + blk] = s;
  }
}

// Final values of a reducing detector at step n: the per-block partials of each value added in
// block order in float64, then
//  energy, flux: the sum as float32, stored in record slot `slot`, or added to it in float32 when
//         the detector accumulates (the value left in the slot is returned to thread 0);
//  field: the box mean of each component (_volume_weighted_spatial_mean, detector.py:95-109,
//         with the uniform cell volume), stored in slot `slot`;
//  phasor: the box mean m of each component, state[f, c] += m * exp(i w_f t_n) * scale * window_n
//          (phasor.py:214-225; the mean of the per-cell products equals this product of the mean).
__device__ __forceinline__ float reduce_finalize_tile(const StepStatic& P, const DetK& K, const double* partial,
                                                      int nparts, void* state, int slot, int n) {
  const int nv = reduced_values(K);
  float first = 0.f;
  if (K.kind == FDTDX_DET_PHASOR && K.cplx) {  // complex fields: the box mean m_r + i m_i of each component
    for (int v = 0; v < K.ncomp; ++v) {
      double ar = 0.0, ai = 0.0;
      for (int b = threadIdx.x; b < nparts; b += kBlock) {
        ar += partial[v * kDetMaxBlocks + b];
        ai += partial[(K.ncomp + v) * kDetMaxBlocks + b];
      }
      const double sr = block_sum(ar);
      const double si = block_sum(ai);
      if (threadIdx.x != 0) continue;
      const float mr = (float)(sr / (double)K.cells), mi = (float)(si / (double)K.cells);
      float2* st = (float2*)state;
      const float w = K.window[n];
      for (int f = 0; f < K.nfreq; ++f) {
        const float cs = K.cos_t[(long long)f * P.T + n], sn = K.sin_t[(long long)f * P.T + n];
        const float re = ((mr * cs) * K.scale) * w - ((mi * sn) * K.scale) * w;
        const float im = ((mi * cs) * K.scale) * w + ((mr * sn) * K.scale) * w;
        float2& o = st[f * K.ncomp + v];
        if (K.sign < 0.f) { o.x = o.x - re; o.y = o.y - im; }
        else { o.x = o.x + re; o.y = o.y + im; }
      }
    }
    return first;
  }
  for (int v = 0; v < kDetMaxValues; ++v) {
    if (v >= nv) break;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) acc += partial[v * kDetMaxBlocks + b];
    const double s = block_sum(acc);
    if (threadIdx.x != 0) continue;
    if (K.kind == FDTDX_DET_ENERGY || K.kind == FDTDX_DET_FLUX) {
      float* o = (float*)state + (long long)slot * nv + v;
      const float r = K.accum ? *o + (float)s : (float)s;
      *o = r;
      if (v == 0) first = r;
    } else {
      const float m = (float)(s / (double)K.cells);
      if (K.kind == FDTDX_DET_FIELD) {
        ((float*)state)[(long long)slot * K.ncomp + v] = m;
      } else {
        float2* st = (float2*)state;
        const float w = K.window[n];
        for (int f = 0; f < K.nfreq; ++f) {
          const float re = ((m * K.cos_t[(long long)f * P.T + n]) * K.scale) * w;
          const float im = ((m * K.sin_t[(long long)f * P.T + n]) * K.scale) * w;
          float2& o = st[f * K.ncomp + v];
          if (K.sign < 0.f) { o.x = o.x - re; o.y = o.y - im; }
          else { o.x = o.x + re; o.y = o.y + im; }
        }
      }
    }
  }
  return first;
}

template <bool CX>
__global__ __launch_bounds__(kBlock) void k_detector(const StepStatic P, const DetK K, const DetDyn D) {
  detector_tile<CX>(P, K, D, D.n, D.slot, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kBlock) void k_reduce_finalize(const StepStatic P, const DetK K, const double* partial,
                                                            int nparts, void* state, int slot, int n) {
  reduce_finalize_tile(P, K, partial, nparts, state, slot, n);
}

static_assert(sizeof(StepStatic) + sizeof(void*) + sizeof(DetGroup) <= 4096, "grouped detector kernel arguments exceed 4 KB");

template <bool CX>
__global__ __launch_bounds__(kBlock) void k_detector_group(const StepStatic P, const DetK* __restrict__ Kd,
                                                           const DetGroup G) {
  int a = 0;
  while (a + 1 < G.count && (int)blockIdx.x >= G.start[a + 1]) ++a;
  const DetK& K = Kd[G.q[a]];
  const DetDyn& D = G.dd[a];
  detector_tile<CX>(P, K, D, D.n, G.slot[a], (int)blockIdx.x - G.start[a], G.start[a + 1] - G.start[a]);
}

__global__ __launch_bounds__(kBlock) void k_finalize_group(const StepStatic P, const DetK* __restrict__ Kd,
                                                           const DetGroup G) {
  const int a0 = G.fin[blockIdx.x];
  for (int a = a0; a < G.count; ++a) {
    const DetK& K = Kd[G.q[a]];
    if (a > a0 && !K.accum) break;
    const DetDyn& D = G.dd[a];
    reduce_finalize_tile(P, K, D.partial, G.start[a + 1] - G.start[a], D.state, G.slot[a], D.n);
    __syncthreads();  // the next accumulating detector reads the value thread 0 just stored
  }
}

// Batch detectors: the blocks of every (entry, detector) pair are concatenated; a persistent grid
// strides over them. Each pair keeps the block count of the single-solve launch, so its partial
// sums, and with them the reduced values, are bit-identical to a solve run alone.
__global__ __launch_bounds__(kBlock) void k_batch_detector(const BatchTable* __restrict__ B,
                                                           const int* __restrict__ base, int r) {
  const int ni = B->ninst;
  const int total = B->det_start[ni];
  int e = 0;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    while (t >= B->det_start[e + 1]) ++e;
    const BatchInst& I = B->inst[e];
    const int n = base[I.slot] + r;
    if (n >= I.stop) continue;
    int b = t - B->det_start[e];
    int q = 0;
    while (q < I.ndet - 1 && b >= I.det_nb[q]) { b -= I.det_nb[q]; ++q; }
    const DetK& K = I.det[q];
    int slot = K.tti[n];
    if (slot < 0) continue;
    if (K.kind == FDTDX_DET_PHASOR) slot = 0;
    detector_tile<false>(I.P, K, I.dd[q], n, slot, b, I.det_nb[q]);
  }
}

__global__ __launch_bounds__(kBlock) void k_batch_finalize(const BatchTable* __restrict__ B,
                                                           const int* __restrict__ base, int r) {
  const int ni = B->ninst;
  const int total = B->fin_start[ni];
  int e = 0;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    while (t >= B->fin_start[e + 1]) ++e;
    const BatchInst& I = B->inst[e];
    const int n = base[I.slot] + r;
    if (n >= I.stop) continue;
    int want = t - B->fin_start[e];
    for (int q = 0; q < I.ndet; ++q) {
      const DetK& K = I.det[q];
      if (!reduced_values(K) || K.accum) continue;
      if (want-- != 0) continue;
      int slot = K.tti[n];
      if (slot >= 0) {
        if (K.kind == FDTDX_DET_PHASOR) slot = 0;
        float v = reduce_finalize_tile(I.P, K, I.dd[q].partial, I.det_nb[q], I.dd[q].state, slot, n);
        // the detectors accumulating into this one's state follow it, in order, in this block
        for (int a = q + 1; a < I.ndet && I.det[a].accum; ++a)
          v = reduce_finalize_tile(I.P, I.det[a], I.dd[a].partial, I.det_nb[a], I.dd[a].state, slot, n);
        if (threadIdx.x == 0 && (K.kind == FDTDX_DET_ENERGY || K.kind == FDTDX_DET_FLUX)) {
          int f = 0;
          if (!isfinite(v)) f = kFlagNonFinite;
          else if (B->blowup > 0.f && fabsf(v) > B->blowup) f = kFlagBlowup;
          if (f) B->flags[I.slot] = B->flags[I.slot] | f;
        }
      }
      break;
    }
  }
}

hipError_t launch_batch_detectors(const BatchTable* tab, const int* base, int r, int grid, bool finalize,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_batch_detector, dim3(grid), dim3(kBlock), 0, s, tab, base, r);
  if (finalize) hipLaunchKernelGGL(k_batch_finalize, dim3(kMaxBatch), dim3(kBlock), 0, s, tab, base, r);
  return hipGetLastError();
}

int detector_blocks(const DetK& K) {
  long long b = (K.cells + kBlock - 1) / kBlock;
  if (b > kDetMaxBlocks) b = kDetMaxBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t launch_detector(const StepStatic& P, const DetK& K, const DetDyn& D, int nblocks, hipStream_t s) {
  if (D.Ei) hipLaunchKernelGGL(k_detector<true>, dim3(nblocks), dim3(kBlock), 0, s, P, K, D);
  else hipLaunchKernelGGL(k_detector<false>, dim3(nblocks), dim3(kBlock), 0, s, P, K, D);
  if (reduced_values(K))
    hipLaunchKernelGGL(k_reduce_finalize, dim3(1), dim3(kBlock), 0, s, P, K, (const double*)D.partial, nblocks,
                       D.state, D.slot, D.n);
  return hipGetLastError();
}

hipError_t launch_detector_group(const StepStatic& P, const DetK* Kdev, const DetGroup& G, hipStream_t s) {
  if (G.count <= 0) return hipSuccess;
  if (G.dd[0].Ei) hipLaunchKernelGGL(k_detector_group<true>, dim3(G.start[G.count]), dim3(kBlock), 0, s, P, Kdev, G);
  else hipLaunchKernelGGL(k_detector_group<false>, dim3(G.start[G.count]), dim3(kBlock), 0, s, P, Kdev, G);
  if (G.nfin > 0) hipLaunchKernelGGL(k_finalize_group, dim3(G.nfin), dim3(kBlock), 0, s, P, Kdev, G);
  return hipGetLastError();
}

// Copy of the three components of a field over box [lo, hi) into dst (3, box); blockIdx.y = 1
// copies the second field (src1 into dst1) over the same box, so that two snapshots share a launch.
__global__ __launch_bounds__(kBlock) void k_snapshot(const StepStatic P, const float* src0, float* dst0,
                                                     const float* src1, float* dst1, int lo0, int lo1, int lo2,
                                                     int hi0, int hi1, int hi2) {
  const float* src = blockIdx.y ? src1 : src0;
  float* dst = blockIdx.y ? dst1 : dst0;
  const int sy = hi1 - lo1, sz = hi2 - lo2;
  const long long cells = (long long)(hi0 - lo0) * sy * sz;
  for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < cells; q += (long long)gridDim.x * kBlock) {
    const long long r = q / sz;
    const int k = (int)(q - r * sz) + lo2;
    const long long ii = r / sy;
    const int j = (int)(r - ii * sy) + lo1;
    const int i = (int)ii + lo0;
    const long long o = ((long long)i * P.ny + j) * P.nz + k;
    for (int c = 0; c < 3; ++c) dst[c * cells + q] = src[c * P.N + o];
  }
}

__global__ __launch_bounds__(kBlock) void k_zero_box(const StepStatic P, float* F, int lo0, int lo1, int lo2,
                                                     int hi0, int hi1, int hi2) {
  const int sy = hi1 - lo1, sz = hi2 - lo2;
  const long long cells = (long long)(hi0 - lo0) * sy * sz;
  for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < cells; q += (long long)gridDim.x * kBlock) {
    const long long r = q / sz;
    const int k = (int)(q - r * sz) + lo2;
    const long long ii = r / sy;
    const int j = (int)(r - ii * sy) + lo1;
    const int i = (int)ii + lo0;
    const long long o = ((long long)i * P.ny + j) * P.nz + k;
    for (int c = 0; c < 3; ++c) F[c * P.N + o] = 0.f;
  }
}

// Zero the three components of F over box [lo, hi) (a vector store per cell and component).
hipError_t launch_zero_box(const StepStatic& P, float* F, const int* lo, const int* hi, hipStream_t s) {
  const long long cells = (long long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
  long long b = (cells + kBlock - 1) / kBlock;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_zero_box, dim3((unsigned)b), dim3(kBlock), 0, s, P, F, lo[0], lo[1], lo[2], hi[0], hi[1],
                     hi[2]);
  return hipGetLastError();
}

hipError_t launch_snapshot(const StepStatic& P, const float* src, float* dst, const int* lo, const int* hi,
                           hipStream_t s, const float* src1, float* dst1) {
  const long long cells = (long long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
  long long b = (cells + kBlock - 1) / kBlock;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_snapshot, dim3((unsigned)b, src1 ? 2u : 1u), dim3(kBlock), 0, s, P, src, dst, src1, dst1,
                     lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
  return hipGetLastError();
}

// ---------------------------------------------------------------- detector adjoint
// Stage 1: cotangents in the interpolated space (6 per region cell), the transpose of each
// detector's update() with respect to its co-located inputs.
__global__ __launch_bounds__(kBlock) void k_det_adj_g(const StepStatic P, const DetK K, const DetAdjDyn D) {
  const Src3 e = box3(D.E1, D.slo, D.shi, 0);
  const Src3 h1 = box3(D.H1, D.slo, D.shi, 1);
  const Src3 h0 = full3(P, D.H0, 1);
  for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < K.cells; q += (long long)gridDim.x * kBlock) {
    int i, j, k;
    region_cell(K, q, i, j, k);
    const long long o = ((long long)i * P.ny + j) * P.nz + k;
    float Ei[3], Hi[3];
    if (K.exact) {
      interp(P, e, h0, h1, i, j, k, Ei, Hi);
    } else {
      for (int c = 0; c < 3; ++c) { Ei[c] = at(P, e, c, i, j, k); Hi[c] = at(P, h1, c, i, j, k); }
    }
    float gE[3] = {0.f, 0.f, 0.f}, gH[3] = {0.f, 0.f, 0.f};
    const float* cf = (const float*)D.cot;
    if (K.kind == FDTDX_DET_FLUX) {  // the face-area weight enters only the reduced sum
      const int na = K.ncomp == 3 ? 3 : 1;
      for (int t = 0; t < na; ++t) {
        const int a = na == 3 ? t : K.axis;
        const float w = K.reduce ? (cf[(long long)D.slot * na + t] * K.weight) * K.sign
                                 : cf[((long long)D.slot * na + t) * K.cells + q] * K.sign;
        const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
        add6(b1, w * pick6(b2 + 3, Ei, Hi), gE, gH);
        add6(b2, -w * pick6(b1 + 3, Ei, Hi), gE, gH);
        add6(b2 + 3, w * pick6(b1, Ei, Hi), gE, gH);
        add6(b1 + 3, -w * pick6(b2, Ei, Hi), gE, gH);
      }
    } else if (K.kind == FDTDX_DET_ENERGY) {
      // u = 1/2 (|E|^2 / inv_eps + |H|^2 / inv_mu); the detector reads inv_eps itself, so its
      // direct derivative -1/2 |E|^2 / inv_eps^2 joins the gradient (one thread per cell)
      float g;
      if (K.reduce == 2) {  // a collapsed-axis mean: the plane cell's cotangent over the axis length
        int b0, b1;
        plane_axes(K.axis, b0, b1);
        const int cc3[3] = {i, j, k};
        const int na = K.hi[K.axis] - K.lo[K.axis];
        const long long q2 = (long long)(cc3[b0] - K.lo[b0]) * (K.hi[b1] - K.lo[b1]) + (cc3[b1] - K.lo[b1]);
        g = cf[(long long)D.slot * (K.cells / na) + q2] / (float)na;
      } else {
        g = K.reduce ? cf[D.slot] * K.weight : cf[(long long)D.slot * K.cells + q];
      }
      float gi = 0.f;
      for (int c = 0; c < 3; ++c) {
        const float iec = D.ie[(D.ec == 3 ? c : 0) * P.N + o];
        gE[c] = g * ((1.0f / iec) * Ei[c]);
        gH[c] = D.mc ? g * ((1.0f / D.im[(D.mc == 3 ? c : 0) * P.N + o]) * Hi[c]) : g * Hi[c];
        const float r = 1.0f / iec;
        const float d = g * ((-0.5f * (r * r)) * (Ei[c] * Ei[c]));
        if (D.ec == 3) D.grad[c * P.N + o] += d;
        else gi += d;
      }
      if (D.ec != 3) D.grad[o] += gi;
    } else if (K.kind == FDTDX_DET_FIELD) {
      for (int c = 0; c < K.ncomp; ++c) {
        const int cc = K.comps[c];
        const float g = K.reduce ? cf[(long long)D.slot * K.ncomp + c] / (float)K.cells
                                 : cf[((long long)D.slot * K.ncomp + c) * K.cells + q];
        add6(cc, g, gE, gH);
      }
    } else {  // cotangent (dL/dRe, dL/dIm) of each complex entry
      const float2* ct = (const float2*)D.cot;
      const float w = K.window[D.n];
      for (int f = 0; f < K.nfreq; ++f) {
        const float cs = K.cos_t[(long long)f * P.T + D.n];
        const float sn = K.sin_t[(long long)f * P.T + D.n];
        for (int c = 0; c < K.ncomp; ++c) {
          const int cc = K.comps[c];
          const float2 g = K.reduce ? ct[f * K.ncomp + c] : ct[((long long)f * K.ncomp + c) * K.cells + q];
          float v = ((g.x * cs + g.y * sn) * K.scale) * w;
          if (K.reduce) v = v / (float)K.cells;
          if (K.sign < 0.f) v = -v;
          add6(cc, v, gE, gH);
        }
      }
    }
    for (int c = 0; c < 3; ++c) {
      D.g[c * K.cells + q] = gE[c];
      D.g[(3 + c) * K.cells + q] = gH[c];
    }
  }
}

__device__ __forceinline__ float gat(const StepStatic& P, const DetK& K, const float* g, int c, int i, int j, int k) {
  // a detector spanning a periodic axis sees its stencil wrap (its other detectors stay one cell
  // away from the faces of periodic axes, so no wrap reaches them)
  if (P.wrap[0]) i = wrap_coord(i, P.nx);
  if (P.wrap[1]) j = wrap_coord(j, P.ny);
  if (P.wrap[2]) k = wrap_coord(k, P.nz);
  if (i < K.lo[0] || j < K.lo[1] || k < K.lo[2] || i >= K.hi[0] || j >= K.hi[1] || k >= K.hi[2]) return 0.f;
  const int sy = K.hi[1] - K.lo[1], sz = K.hi[2] - K.lo[2];
  return g[c * K.cells + ((long long)(i - K.lo[0]) * sy + (j - K.lo[1])) * sz + (k - K.lo[2])];
}

// Transposed interpolation stencils (exact detectors) gathered at cell (i, j, k): vE from the E
// stencils, vH from the time-centred H ones (the caller splits vH between H^n and H^{n+1}).
__device__ __forceinline__ void gather_T(const StepStatic& P, const DetK& K, const float* g, int i, int j, int k,
                                         float* vE, float* vH) {
  vE[0] = 0.25f * (gat(P, K, g, 0, i, j, k) + gat(P, K, g, 0, i + 1, j, k) + gat(P, K, g, 0, i, j, k - 1) +
                   gat(P, K, g, 0, i + 1, j, k - 1));
  vE[1] = 0.25f * (gat(P, K, g, 1, i, j, k) + gat(P, K, g, 1, i, j + 1, k) + gat(P, K, g, 1, i, j, k - 1) +
                   gat(P, K, g, 1, i, j + 1, k - 1));
  vE[2] = gat(P, K, g, 2, i, j, k);
  vH[0] = 0.5f * (gat(P, K, g, 3, i, j, k) + gat(P, K, g, 3, i, j + 1, k));
  vH[1] = 0.5f * (gat(P, K, g, 4, i, j, k) + gat(P, K, g, 4, i + 1, j, k));
  float s = 0.f;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) s += gat(P, K, g, 5, i + a, j + b, k - c);
  vH[2] = 0.125f * s;
}

// Stage 2: gather the transposed interpolation stencils onto the expanded region and add them to
// the adjoint fields: lam_E += v_E, lam_H += v_Hc / 2; the other half belongs to H^n and is added
// after the adjoint sweeps (k_add_half_vH). Below an electric symmetry plane the stencils read the
// mirror of rows 0 and 1 (mirror_fold), so the cotangent gathered at each halo position folds back
// onto its mirror row with the component's sign, for every set of mirror axes the cell borders.
__global__ __launch_bounds__(kBlock) void k_det_adj_scatter(const StepStatic P, const DetK K, const DetAdjDyn D) {
  const int ey = D.ehi[1] - D.elo[1], ez = D.ehi[2] - D.elo[2];
  const long long cells = (long long)(D.ehi[0] - D.elo[0]) * ey * ez;
  for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < cells; q += (long long)gridDim.x * kBlock) {
    const long long r = q / ez;
    const int k = (int)(q - r * ez) + D.elo[2];
    const long long ii = r / ey;
    const int j = (int)(r - ii * ey) + D.elo[1];
    const int i = (int)ii + D.elo[0];
    const long long o = ((long long)i * P.ny + j) * P.nz + k;
    float vE[3], vH[3];
    if (K.exact) {
      gather_T(P, K, D.g, i, j, k, vE, vH);
      const int cc[3] = {i, j, k};
      int near = 0;  // mirror axes where this cell is row 0 or 1
      for (int a = 0; a < 3; ++a)
        if (P.mirror[a] && cc[a] <= 1) near |= 1 << a;
      for (int S = near; S; S = (S - 1) & near) {  // every nonempty subset of them
        int hq[3] = {i, j, k};
        for (int a = 0; a < 3; ++a)
          if ((S >> a) & 1) hq[a] = -1;
        float hE[3], hH[3];
        gather_T(P, K, D.g, hq[0], hq[1], hq[2], hE, hH);
        for (int c = 0; c < 3; ++c) {
          float sE = 1.f, sH = 1.f;
          bool tE = true, tH = true;
          for (int a = 0; a < 3; ++a) {
            if (!((S >> a) & 1)) continue;
            const bool onE = c != a, onH = c == a;  // sampled on the plane: odd, mirrors row 1
            tE = tE && cc[a] == (onE ? 1 : 0);
            tH = tH && cc[a] == (onH ? 1 : 0);
            if (onE) sE = -sE;
            if (onH) sH = -sH;
          }
          if (tE) vE[c] += sE * hE[c];
          if (tH) vH[c] += sH * hH[c];
        }
      }
      for (int c = 0; c < 3; ++c) {
        D.lamE[c * P.N + o] += vE[c];
        const float hv = 0.5f * vH[c];
        D.lamH[c * P.N + o] += hv;
        D.half_vH[c * cells + q] = hv;
      }
    } else {
      for (int c = 0; c < 3; ++c) {
        D.lamE[c * P.N + o] += gat(P, K, D.g, c, i, j, k);
        D.lamH[c * P.N + o] += gat(P, K, D.g, 3 + c, i, j, k);
        D.half_vH[c * cells + q] = 0.f;
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_add_half_vH(const StepStatic P, const DetAdjDyn D) {
  const int ey = D.ehi[1] - D.elo[1], ez = D.ehi[2] - D.elo[2];
  const long long cells = (long long)(D.ehi[0] - D.elo[0]) * ey * ez;
  for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < cells; q += (long long)gridDim.x * kBlock) {
    const long long r = q / ez;
    const int k = (int)(q - r * ez) + D.elo[2];
    const long long ii = r / ey;
    const int j = (int)(r - ii * ey) + D.elo[1];
    const int i = (int)ii + D.elo[0];
    const long long o = ((long long)i * P.ny + j) * P.nz + k;
    for (int c = 0; c < 3; ++c) D.lamH[c * P.N + o] += D.half_vH[c * cells + q];
  }
}

static int blocks_for(long long cells) {
  long long b = (cells + kBlock - 1) / kBlock;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t launch_detector_adjoint(const StepStatic& P, const DetK& K, const DetAdjDyn& D, hipStream_t s) {
  hipLaunchKernelGGL(k_det_adj_g, dim3(blocks_for(K.cells)), dim3(kBlock), 0, s, P, K, D);
  const long long ec = (long long)(D.ehi[0] - D.elo[0]) * (D.ehi[1] - D.elo[1]) * (D.ehi[2] - D.elo[2]);
  hipLaunchKernelGGL(k_det_adj_scatter, dim3(blocks_for(ec)), dim3(kBlock), 0, s, P, K, D);
  return hipGetLastError();
}

hipError_t launch_add_half_vH(const StepStatic& P, const DetAdjDyn& D, hipStream_t s) {
  const long long ec = (long long)(D.ehi[0] - D.elo[0]) * (D.ehi[1] - D.elo[1]) * (D.ehi[2] - D.elo[2]);
  hipLaunchKernelGGL(k_add_half_vH, dim3(blocks_for(ec)), dim3(kBlock), 0, s, P, D);
  return hipGetLastError();
}

}  // namespace fdtdx
