// Register-march single-pass Yee step: one kernel computes E^{n+1} and H^{n+1}, reading
// (E^n, H^n, inv_eps, psi^n) and writing (E^{n+1}, H^{n+1}, psi^{n+1}) once per cell, with no LDS
// and no barriers. The float32 operations and their order are those of the two-sweep kernels
// (curl.py:265-279, 357-366; PML.py:138-190; update.py:298-354, 736-750), so the results are
// bit-identical to them.
//
// Work decomposition. Each wavefront is independent and owns a tile: a run of W consecutive z
// cells (one per lane, z contiguous so every row access is one coalesced 256 B transaction),
// TY rows of y (held in registers) and a chunk of x-planes it marches through. At plane i the
// wave
//   - computes E^{n+1}(i) row by row for rows j0 .. j0+TY (the last row is a recomputed halo)
//     from H^n(i) (loaded), H^n(i-1) (registers, previous plane), H^n(i, j-1) (previous row) and
//     H^n(i, j, k-1) (the neighbour lane, by a lane shuffle);
//   - finishes H^{n+1}(i-1) for its owned rows: the y/z derivatives of E^{n+1}(i-1) were taken one
//     plane earlier and kept in registers, the x derivative needs E^{n+1}(i) just computed.
// Lane W computes E^{n+1} of the first z cell past the tile (halo, not stored) for the z
// derivative of the last owned lane; lane 0 loads H^n(k-1) itself. Neighbouring tiles recompute
// each other's halo cells, so every state array is ping-ponged (read the "in" copy, write the
// "out" copy) and the ADE polarisation uses three rotating slots.
#include "engine_common.h"
#include "kernels.h"
#include "stencil.h"

namespace fdtdx {

// Loads and stores at a wave-uniform base plus a 32-bit byte offset, the form the global
// instructions take with a scalar base register (no 64-bit address per lane).
__device__ __forceinline__ float ldb(const float* base, unsigned ob) {
  return *(const float*)((const char*)base + ob);
}
__device__ __forceinline__ void stb(float* base, unsigned ob, float v) { *(float*)((char*)base + ob) = v; }

// Inputs of one x-plane for a wavefront's rows: H^n of row j0-1 (x, z), H^n, E^n and inv_eps of
// rows j0 .. j0+TY, and for lane 0 H^n (x, y) of the z cell before the tile.
template <int EC, int TY>
struct MarchPlane {
  float hjm[2];
  float h[TY + 1][3];
  float e[TY + 1][3];
  float ie[TY + 1][EC];
  float hk[TY + 1][2];
};

template <int EC, int TY>
__device__ __forceinline__ void march_load(MarchPlane<EC, TY>& L, int i, int Nx, int Ny, int Nz, int j0, int k, int lane,
                                           bool act, const float* Hx, const float* Hy, const float* Hz,
                                           const float* Ex, const float* Ey, const float* Ez, const float* iex,
                                           const float* iey, const float* iez) {
  L.hjm[0] = L.hjm[1] = 0.f;
#pragma unroll
  for (int r = 0; r <= TY; ++r) {
    L.h[r][0] = L.h[r][1] = L.h[r][2] = 0.f;
    L.e[r][0] = L.e[r][1] = L.e[r][2] = 0.f;
#pragma unroll
    for (int q = 0; q < EC; ++q) L.ie[r][q] = 0.f;
    L.hk[r][0] = L.hk[r][1] = 0.f;
  }
  if (i >= Nx || !act) return;
  if (j0 > 0) {
    const unsigned ob = (unsigned)((i * Ny + (j0 - 1)) * Nz + k) * 4u;
    L.hjm[0] = ldb(Hx, ob);
    L.hjm[1] = ldb(Hz, ob);
  }
#pragma unroll
  for (int r = 0; r <= TY; ++r) {
    if (j0 + r < Ny) {
      const unsigned ob = (unsigned)((i * Ny + (j0 + r)) * Nz + k) * 4u;
      L.h[r][0] = ldb(Hx, ob);
      L.h[r][1] = ldb(Hy, ob);
      L.h[r][2] = ldb(Hz, ob);
      L.e[r][0] = ldb(Ex, ob);
      L.e[r][1] = ldb(Ey, ob);
      L.e[r][2] = ldb(Ez, ob);
      L.ie[r][0] = ldb(iex, ob);
      if constexpr (EC == 3) {
        L.ie[r][1] = ldb(iey, ob);
        L.ie[r][2] = ldb(iez, ob);
      }
      if (lane == 0 && k > 0) {
        L.hk[r][0] = ldb(Hx, ob - 4u);
        L.hk[r][1] = ldb(Hy, ob - 4u);
      }
    }
  }
}

__device__ __forceinline__ bool span_hits(int a0, int a1, const int lo, const int hi) {  // [a0, a1] vs [lo, hi)
  return a1 >= lo && a0 < hi;
}

template <int EC, int MC, int TY>
__global__ __launch_bounds__(kMarchThreads) void k_march(const StepStatic P, const MarchDyn D) {
  const int lane = threadIdx.x & 63;
  const long long lb = logical_block(0, D.nblocks_pad);
  const long long t = lb * (kMarchThreads / 64) + (threadIdx.x >> 6);
  if (lb >= D.nblocks || t >= D.ntiles) return;  // whole wavefronts leave; there are no barriers
  const int zc = (int)(t % D.nzc);
  const long long yx = t / D.nzc;
  const int yt = (int)(yx % D.nyt);
  const int xc = (int)(yx / D.nyt);
  const int Nx = P.nx, Ny = P.ny, Nz = P.nz;
  const long long N = P.N;
  const float c = P.c;
  const int n = D.n;
  const int kz0 = zc * D.W;
  if (kz0 >= Nz) return;
  const int wz = min(D.W, Nz - kz0);
  const int k = kz0 + lane;
  const bool own_k = lane < wz;
  const bool act = lane <= wz && k < Nz;  // computes E^{n+1}: owned lanes plus the halo lane
  const int j0 = yt * TY;
  const int x0 = xc * D.XC;
  const int x1 = min(x0 + D.XC, Nx);
  // Component base pointers are wave-uniform; cell offsets are 32-bit (the plan guarantees
  // 3 * N < 2^31), so every hot access is a scalar base plus a 32-bit byte offset.
  const float* __restrict__ Hx_in = D.H_in;
  const float* __restrict__ Hy_in = D.H_in + N;
  const float* __restrict__ Hz_in = D.H_in + 2 * N;
  const float* __restrict__ Ex_in = D.E_in;
  const float* __restrict__ Ey_in = D.E_in + N;
  const float* __restrict__ Ez_in = D.E_in + 2 * N;
  float* __restrict__ Ex_out = D.E_out;
  float* __restrict__ Ey_out = D.E_out + N;
  float* __restrict__ Ez_out = D.E_out + 2 * N;
  float* __restrict__ Hx_out = D.H_out;
  float* __restrict__ Hy_out = D.H_out + N;
  float* __restrict__ Hz_out = D.H_out + 2 * N;
  const float* __restrict__ iex_p = D.ie;
  const float* __restrict__ iey_p = D.ie + (EC == 3 ? N : 0);
  const float* __restrict__ iez_p = D.ie + (EC == 3 ? 2 * N : 0);
  // wave footprint across y and z (rows j0 .. j0+TY, lanes kz0 .. kz0+wz)
  const int ylo = j0, yhi = j0 + TY, zlo = kz0, zhi = kz0 + wz;

  float hp[TY + 1][3];  // H^n(i-1, j0+r)
  float dv[TY][4];      // of E^{n+1}(i-1, j0+r): dyEz, dzEy, dzEx, dyEx
  float ep[TY][3];      // E^{n+1}(i-1, j0+r)
#pragma unroll
  for (int r = 0; r <= TY; ++r) {
    hp[r][0] = hp[r][1] = hp[r][2] = 0.f;
    if (x0 > 0 && act && j0 + r < Ny) {
      const unsigned ob = (unsigned)(((x0 - 1) * Ny + (j0 + r)) * Nz + k) * 4u;
      hp[r][0] = ldb(Hx_in, ob);
      hp[r][1] = ldb(Hy_in, ob);
      hp[r][2] = ldb(Hz_in, ob);
    }
  }
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    dv[r][0] = dv[r][1] = dv[r][2] = dv[r][3] = 0.f;
    ep[r][0] = ep[r][1] = ep[r][2] = 0.f;
  }

  MarchPlane<EC, TY> cur, nxt;
  march_load<EC, TY>(cur, x0, Nx, Ny, Nz, j0, k, lane, act, Hx_in, Hy_in, Hz_in, Ex_in, Ey_in, Ez_in, iex_p, iey_p,
                     iez_p);
  for (int i = x0; i <= x1; ++i) {
    // prefetch the next plane; its loads are in flight while this plane is computed
    if (i < x1)
      march_load<EC, TY>(nxt, i + 1, Nx, Ny, Nz, j0, k, lane, act, Hx_in, Hy_in, Hz_in, Ex_in, Ey_in, Ez_in, iex_p,
                         iey_p, iez_p);
    const bool doE = i < Nx;        // E^{n+1}(Nx) is the zero halo
    const bool ownE = i < x1;       // plane x1 is recomputed for the x derivative only
    const bool doH = i > x0;        // H^{n+1}(i-1) is owned from the second plane on
    // "plain": nothing but the bare curl updates touch this plane's footprint (planes i-1 and i)
    bool plain = true;
    for (int s = 0; s < P.nslab; ++s) {
      const SlabK& sl = P.slab[s];
      if (sl.axis == 0) plain = plain && !span_hits(i - 1, i, sl.lo, sl.hi);
      else if (sl.axis == 1) plain = plain && !span_hits(ylo, yhi, sl.lo, sl.hi);
      else plain = plain && !span_hits(zlo, zhi, sl.lo, sl.hi);
    }
    for (int s = 0; s < P.nsrc; ++s) {
      const SrcK& sr = P.src[s];
      plain = plain && !(span_hits(i - 1, i, sr.lo[0], sr.hi[0]) && span_hits(ylo, yhi, sr.lo[1], sr.hi[1]) &&
                         span_hits(zlo, zhi, sr.lo[2], sr.hi[2]));
    }
    for (int b = 0; b < P.nade; ++b) {
      const AdeK& A = P.ade[b];
      plain = plain && !(span_hits(i, i, A.lo[0], A.hi[0]) && span_hits(ylo, yhi, A.lo[1], A.hi[1]) &&
                         span_hits(zlo, zhi, A.lo[2], A.hi[2]));
    }
    if (D.snap)
      plain = plain && !(span_hits(i - 1, i - 1, D.slo[0], D.shi[0]) && span_hits(ylo, yhi, D.slo[1], D.shi[1]) &&
                         span_hits(zlo, zhi, D.slo[2], D.shi[2]));

    float ex_pr = 0.f, ey_pr = 0.f, ez_pr = 0.f, dzEy_pr = 0.f, dzEx_pr = 0.f;  // E^{n+1}(i, j-1)
#pragma unroll
    for (int r = 0; r <= TY; ++r) {
      const int j = j0 + r;
      const bool rowv = j < Ny;
      const bool own_row = r < TY && rowv;
      float exn = 0.f, eyn = 0.f, ezn = 0.f;
      const float hx = cur.h[r][0], hy = cur.h[r][1], hz = cur.h[r][2];
      if (doE && rowv) {  // wave-uniform
        const float hx_jm = r == 0 ? cur.hjm[0] : cur.h[r - 1][0];
        const float hz_jm = r == 0 ? cur.hjm[1] : cur.h[r - 1][2];
        float hx_km = __shfl_up(hx, 1);
        float hy_km = __shfl_up(hy, 1);
        if (lane == 0) {
          hx_km = cur.hk[r][0];
          hy_km = cur.hk[r][1];
        }
        if (act) {
          const int o = (i * Ny + j) * Nz + k;
          const unsigned ob = (unsigned)o * 4u;
          Deriv d;
          d.dy_z = hz - hz_jm; d.dz_y = hy - hy_km; d.dz_x = hx - hx_km;
          d.dx_z = hz - hp[r][2]; d.dx_y = hy - hp[r][1]; d.dy_x = hx - hx_jm;
          d.dx_x = d.dy_y = d.dz_z = 0.f;
          float cx = d.dy_z - d.dz_y;
          float cy = d.dz_x - d.dx_z;
          float cz = d.dx_y - d.dy_x;
          const bool st = own_k && own_row && ownE;
          if (!plain) apply_cpml<false>(P, kMaskAll, D.psi_in, D.psi_out, 0, st ? 1 : 2, i, j, k, d, cx, cy, cz);
          const float iex = cur.ie[r][0];
          const float iey = EC == 3 ? cur.ie[r][EC == 3 ? 1 : 0] : iex;
          const float iez = EC == 3 ? cur.ie[r][EC == 3 ? 2 : 0] : iex;
          const float ex = cur.e[r][0], ey = cur.e[r][1], ez = cur.e[r][2];
          exn = ex + (c * cx) * iex;
          eyn = ey + (c * cy) * iey;
          ezn = ez + (c * cz) * iez;
          if (!plain) {
#pragma unroll 1
            for (int b = 0; b < FDTDX_MAX_ADE; ++b) {
              if (b >= P.nade) break;
              const AdeK& A = P.ade[b];
              if (!in_box(i, j, k, A.lo, A.hi)) continue;
              const long long aq = ((long long)(i - A.lo[0]) * A.by + (j - A.lo[1])) * A.bz + (k - A.lo[2]);
              const long long slot = (long long)A.poles * 3 * A.cells;
              const float* pcur = D.ade + A.state_off + (long long)(n % 3) * slot;
              const float* pprv = D.ade + A.state_off + (long long)((n + 2) % 3) * slot;
              float* pnxt = D.ade + A.state_off + (long long)((n + 1) % 3) * slot;
              float dl[3] = {0.f, 0.f, 0.f};
              const float eo[3] = {ex, ey, ez};
#pragma unroll
              for (int comp = 0; comp < 3; ++comp) {
                for (int p = 0; p < A.poles; ++p) {
                  const int cc = A.ccomp == 3 ? comp : 0;
                  const long long ci = ((long long)p * A.ccomp + cc) * A.cells + aq;
                  const float c1 = A.coef[ci];
                  const float c2 = A.coef[(long long)A.poles * A.ccomp * A.cells + ci];
                  const float c3 = A.coef[2ll * A.poles * A.ccomp * A.cells + ci];
                  const long long si = ((long long)p * 3 + comp) * A.cells + aq;
                  const float pc = pcur[si], pp = pprv[si];
                  const float ph = (c1 * pc + c2 * pp) + c3 * eo[comp];
                  dl[comp] = p == 0 ? pc - ph : dl[comp] + (pc - ph);
                  if (st) pnxt[si] = ph;
                }
              }
              exn = exn + iex * dl[0];
              eyn = eyn + iey * dl[1];
              ezn = ezn + iez * dl[2];
            }
            apply_sources<false>(P, kMaskAll, D.ie, D.im, 0, n, false, i, j, k, o, EC, 0, exn, eyn, ezn);
          }
          if (st) {
            stb(Ex_out, ob, exn);
            stb(Ey_out, ob, eyn);
            stb(Ez_out, ob, ezn);
          }
        }
      }
      // z derivatives of E^{n+1}(i, j) need the next lane's values (zero past the domain)
      float ex_kp = __shfl_down(exn, 1);
      float ey_kp = __shfl_down(eyn, 1);
      if (k + 1 >= Nz) {
        ex_kp = 0.f;
        ey_kp = 0.f;
      }
      // H^{n+1}(i-1, j) (update.py:736-750) for owned rows and lanes
      if (doH && own_row && own_k) {
        const int ip = i - 1;
        Deriv d;
        d.dy_z = dv[r][0]; d.dz_y = dv[r][1]; d.dz_x = dv[r][2]; d.dy_x = dv[r][3];
        d.dx_z = ezn - ep[r][2];
        d.dx_y = eyn - ep[r][1];
        d.dx_x = d.dy_y = d.dz_z = 0.f;
        float cx = d.dy_z - d.dz_y;
        float cy = d.dz_x - d.dx_z;
        float cz = d.dx_y - d.dy_x;
        if (!plain) apply_cpml<false>(P, kMaskAll, D.psi_in, D.psi_out, 1, 1, ip, j, k, d, cx, cy, cz);
        const int o = (ip * Ny + j) * Nz + k;
        const unsigned ob = (unsigned)o * 4u;
        const float imx = MC ? D.im[o] : 1.0f;
        const float imy = MC == 3 ? D.im[N + o] : imx;
        const float imz = MC == 3 ? D.im[2 * N + o] : imx;
        float hxn = hp[r][0] - (c * cx) * imx;
        float hyn = hp[r][1] - (c * cy) * imy;
        float hzn = hp[r][2] - (c * cz) * imz;
        if (!plain) apply_sources<false>(P, kMaskAll, D.ie, D.im, 1, n, false, ip, j, k, o, 1, MC, hxn, hyn, hzn);
        stb(Hx_out, ob, hxn);
        stb(Hy_out, ob, hyn);
        stb(Hz_out, ob, hzn);
        if (!plain && D.snap && in_box(ip, j, k, D.slo, D.shi)) {
          const long long sy = D.shi[1] - D.slo[1], sz = D.shi[2] - D.slo[2];
          const long long sc = (long long)(D.shi[0] - D.slo[0]) * sy * sz;
          const long long sq = ((long long)(ip - D.slo[0]) * sy + (j - D.slo[1])) * sz + (k - D.slo[2]);
          D.snap[sq] = hp[r][0];
          D.snap[sc + sq] = hp[r][1];
          D.snap[2 * sc + sq] = hp[r][2];
        }
        if (!plain && D.record) {  // collect_interfaces (update.py:1140-1178)
#pragma unroll 1
          for (int s = 0; s < FDTDX_MAX_CPML; ++s) {
            if (s >= P.nslab) break;
            const SlabK& sl = P.slab[s];
            const int coord = sl.axis == 0 ? ip : (sl.axis == 1 ? j : k);
            if (coord != sl.iface) continue;
            const long long pc = (long long)sl.pn0 * sl.pn1;
            const long long rq = sl.axis == 0 ? (long long)j * Nz + k
                                 : (sl.axis == 1 ? (long long)ip * Nz + k : (long long)ip * Ny + j);
            float* re = (float*)(D.record + sl.rec_off_E);  // float32 layers only
            float* rh = (float*)(D.record + sl.rec_off_H);
            re[rq] = ep[r][0];
            re[pc + rq] = ep[r][1];
            re[2 * pc + rq] = ep[r][2];
            rh[rq] = hxn;
            rh[pc + rq] = hyn;
            rh[2 * pc + rq] = hzn;
          }
        }
      }
      // row r-1 of plane i is complete now that row r is known: its y derivatives
      if (r >= 1) {
        dv[r - 1][0] = ezn - ez_pr;
        dv[r - 1][1] = dzEy_pr;
        dv[r - 1][2] = dzEx_pr;
        dv[r - 1][3] = exn - ex_pr;
        ep[r - 1][0] = ex_pr;
        ep[r - 1][1] = ey_pr;
        ep[r - 1][2] = ez_pr;
      }
      hp[r][0] = hx;
      hp[r][1] = hy;
      hp[r][2] = hz;
      ex_pr = exn;
      ey_pr = eyn;
      ez_pr = ezn;
      dzEy_pr = ey_kp - eyn;
      dzEx_pr = ex_kp - exn;
    }
    cur = nxt;
  }
}

// Tile shape: W lanes of z (the z extent split into near-equal runs of at most 63 so the halo
// lane fits the wavefront), TY rows of y, XC planes of x per wavefront.
bool march_config(const StepStatic& P, int ty, int xc, MarchDyn& D) {
  if (ty != 2 && ty != 4) return false;
  if (P.nz < 1 || P.ny < 1 || P.nx < 1) return false;
  if (3 * P.N >= (1ll << 31)) return false;  // 32-bit cell offsets
  D.nzc = (P.nz + 62) / 63;
  D.W = (P.nz + D.nzc - 1) / D.nzc;
  D.TY = ty;
  D.nyt = (P.ny + ty - 1) / ty;
  if (xc < 1) xc = 1;
  if (xc > P.nx) xc = P.nx;
  D.XC = xc;
  D.nxc = (P.nx + xc - 1) / xc;
  D.ntiles = (long long)D.nzc * D.nyt * D.nxc;
  D.nblocks = (D.ntiles + (kMarchThreads / 64) - 1) / (kMarchThreads / 64);
  D.nblocks_pad = pad_blocks(D.nblocks);
  return true;
}

template <int TY>
static void launch_march_ty(const StepStatic& P, const MarchDyn& D, int ec, int mc, hipStream_t s) {
  const dim3 g((unsigned)D.nblocks_pad), b(kMarchThreads);
  if (ec == 3) {
    if (mc == 3) hipLaunchKernelGGL((k_march<3, 3, TY>), g, b, 0, s, P, D);
    else if (mc == 1) hipLaunchKernelGGL((k_march<3, 1, TY>), g, b, 0, s, P, D);
    else hipLaunchKernelGGL((k_march<3, 0, TY>), g, b, 0, s, P, D);
  } else {
    if (mc == 3) hipLaunchKernelGGL((k_march<1, 3, TY>), g, b, 0, s, P, D);
    else if (mc == 1) hipLaunchKernelGGL((k_march<1, 1, TY>), g, b, 0, s, P, D);
    else hipLaunchKernelGGL((k_march<1, 0, TY>), g, b, 0, s, P, D);
  }
}

hipError_t launch_march(const StepStatic& P, const MarchDyn& D, int ec, int mc, hipStream_t s) {
  if (D.TY == 2) launch_march_ty<2>(P, D, ec, mc, s);
  else launch_march_ty<4>(P, D, ec, mc, s);
  return hipGetLastError();
}

}  // namespace fdtdx
