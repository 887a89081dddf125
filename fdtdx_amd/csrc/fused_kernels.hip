// Fused single-pass Yee step: one kernel computes E^{n+1} and H^{n+1} for a tile, reading
// (E^n, H^n, inv_eps, psi^n) once and writing (E^{n+1}, H^{n+1}, psi^{n+1}) once, i.e. 13 words per
// cell instead of the 19 of separate E and H sweeps. Same float32 operations in the same order as
// the two-sweep kernels (curl.py, PML.py, update.py), so results are bit-identical to them.
//
// Work decomposition: a workgroup owns TY rows (y) x all of z for a chunk of x-planes and marches
// along x. Per plane i it
//   P1  stages H^n(i) for rows [j0-1, j1] in LDS,
//   P2  computes E^{n+1}(i) for rows [j0, j1] (row j1 is a recomputed halo, not stored) and stages
//       it in LDS,
//   P3  finishes H^{n+1}(i-1) for the owned rows: its y/z derivatives of E^{n+1}(i-1) were taken in
//       the previous iteration and kept in registers, the x derivative needs E^{n+1}(i) at the same
//       (j, k), which this thread just computed.
// Because neighbouring tiles recompute each other's halo, every state array is ping-ponged (read
// the "in" copy, write the "out" copy) and the ADE polarisation uses three rotating slots.
#include "engine_common.h"
#include "kernels.h"
#include "stencil.h"

namespace fdtdx {

template <int EC, int MC, int M>
__global__ __launch_bounds__(kFusedThreads) void k_fused(const StepStatic P, const FusedDyn D) {
  extern __shared__ float lds[];
  const long long lb = logical_block(0, D.nblocks_pad);
  if (lb >= D.nblocks) return;
  const int tile = (int)(lb % D.ntiles);
  const int chunk = (int)(lb / D.ntiles);
  const int Nx = P.nx, Ny = P.ny, Nz = P.nz;
  const int j0 = tile * D.TY;
  const int j1 = min(j0 + D.TY, Ny);
  const int x0 = chunk * D.chunk;
  const int x1 = min(x0 + D.chunk, Nx);
  const int rowsE = j1 - j0 + 1;
  const int rowsH = j1 - j0 + 2;
  const int strideH = (D.TY + 2) * Nz;
  const int strideE = (D.TY + 1) * Nz;
  float* Hs = lds;
  float* Es = lds + 3 * strideH;
  const int tid = threadIdx.x;
  const long long N = P.N;
  const float c = P.c;

  int pj[M], pk[M];
  bool pv[M], po[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int q = tid + kFusedThreads * m;
    const int jj = q / Nz;
    pk[m] = q - jj * Nz;
    pj[m] = j0 + jj;
    pv[m] = q < rowsE * Nz && pj[m] < Ny;
    po[m] = pv[m] && pj[m] < j1;
  }
  float hp[M][3], ep[M][2], dv[M][4], en[M][2];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    hp[m][0] = hp[m][1] = hp[m][2] = 0.f;
    ep[m][0] = ep[m][1] = 0.f;
    dv[m][0] = dv[m][1] = dv[m][2] = dv[m][3] = 0.f;
    if (pv[m] && x0 > 0) {
      const long long o = ((long long)(x0 - 1) * Ny + pj[m]) * Nz + pk[m];
      hp[m][0] = D.H_in[o];
      hp[m][1] = D.H_in[N + o];
      hp[m][2] = D.H_in[2 * N + o];
    }
  }

  for (int i = x0; i <= x1; ++i) {
    const bool doE = i < Nx;
    // ---- P1: H^n(i) rows [j0-1, j1] -> LDS
    for (int q = tid; q < rowsH * Nz; q += kFusedThreads) {
      const int jj = q / Nz;
      const int k = q - jj * Nz;
      const int j = j0 - 1 + jj;
      float vx = 0.f, vy = 0.f, vz = 0.f;
      if (doE && j >= 0 && j < Ny) {
        const long long o = ((long long)i * Ny + j) * Nz + k;
        vx = D.H_in[o];
        vy = D.H_in[N + o];
        vz = D.H_in[2 * N + o];
      }
      Hs[q] = vx;
      Hs[strideH + q] = vy;
      Hs[2 * strideH + q] = vz;
    }
    __syncthreads();
    // ---- P2: E^{n+1}(i) on rows [j0, j1]
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int q = tid + kFusedThreads * m;
      float ex_n = 0.f, ey_n = 0.f, ez_n = 0.f;
      if (pv[m] && doE) {
        const int j = pj[m], k = pk[m];
        const int jh = (j - j0 + 1) * Nz + k;  // row of (j, k) in Hs
        const float hx = Hs[jh], hy = Hs[strideH + jh], hz = Hs[2 * strideH + jh];
        const float hx_jm = Hs[jh - Nz], hz_jm = Hs[2 * strideH + jh - Nz];
        const float hx_km = k > 0 ? Hs[jh - 1] : 0.f;
        const float hy_km = k > 0 ? Hs[strideH + jh - 1] : 0.f;
        Deriv d;
        d.dy_z = hz - hz_jm; d.dz_y = hy - hy_km; d.dz_x = hx - hx_km;
        d.dx_z = hz - hp[m][2]; d.dx_y = hy - hp[m][1]; d.dy_x = hx - hx_jm;
        d.dx_x = d.dy_y = d.dz_z = 0.f;
        float cx = d.dy_z - d.dz_y;
        float cy = d.dz_x - d.dx_z;
        float cz = d.dx_y - d.dy_x;
        const bool store = po[m] && i < x1;
        apply_cpml<false>(P, kMaskAll, D.psi_in, D.psi_out, 0, store ? 1 : 2, i, j, k, d, cx, cy, cz);
        const long long o = ((long long)i * Ny + j) * Nz + k;
        const float iex = D.ie[o];
        const float iey = EC == 3 ? D.ie[N + o] : iex;
        const float iez = EC == 3 ? D.ie[2 * N + o] : iex;
        const float ex = D.E_in[o], ey = D.E_in[N + o], ez = D.E_in[2 * N + o];
        ex_n = ex + (c * cx) * iex;
        ey_n = ey + (c * cy) * iey;
        ez_n = ez + (c * cz) * iez;
#pragma unroll 1
        for (int b = 0; b < FDTDX_MAX_ADE; ++b) {
          if (b >= P.nade) break;
          const AdeK& A = P.ade[b];
          if (!in_box(i, j, k, A.lo, A.hi)) continue;
          const long long aq = ((long long)(i - A.lo[0]) * A.by + (j - A.lo[1])) * A.bz + (k - A.lo[2]);
          const long long slot = (long long)A.poles * 3 * A.cells;
          const float* cur = D.ade + A.state_off + (long long)(D.n % 3) * slot;
          const float* prv = D.ade + A.state_off + (long long)((D.n + 2) % 3) * slot;
          float* nxt = D.ade + A.state_off + (long long)((D.n + 1) % 3) * slot;
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
              const float pc = cur[si], pp = prv[si];
              const float ph = (c1 * pc + c2 * pp) + c3 * eo[comp];
              dl[comp] = p == 0 ? pc - ph : dl[comp] + (pc - ph);
              if (store) nxt[si] = ph;
            }
          }
          ex_n = ex_n + iex * dl[0];
          ey_n = ey_n + iey * dl[1];
          ez_n = ez_n + iez * dl[2];
        }
        apply_sources<false>(P, kMaskAll, D.ie, D.im, 0, D.n, false, i, j, k, o, EC, 0, ex_n, ey_n, ez_n);
        if (store) {
          D.E_out[o] = ex_n;
          D.E_out[N + o] = ey_n;
          D.E_out[2 * N + o] = ez_n;
        }
      }
      if (q < rowsE * Nz) {
        Es[q] = ex_n;
        Es[strideE + q] = ey_n;
        Es[2 * strideE + q] = ez_n;
      }
      en[m][0] = ey_n;
      en[m][1] = ez_n;
    }
    __syncthreads();
    // ---- P3: H^{n+1}(i-1) on owned rows; y/z derivatives of E^{n+1}(i) for the next plane
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int j = pj[m], k = pk[m];
      if (po[m] && i > x0) {
        const int ip = i - 1;
        Deriv d;
        d.dy_z = dv[m][0]; d.dz_y = dv[m][1]; d.dz_x = dv[m][2]; d.dy_x = dv[m][3];
        d.dx_z = en[m][1] - ep[m][1];
        d.dx_y = en[m][0] - ep[m][0];
        d.dx_x = d.dy_y = d.dz_z = 0.f;
        float cx = d.dy_z - d.dz_y;
        float cy = d.dz_x - d.dx_z;
        float cz = d.dx_y - d.dy_x;
        apply_cpml<false>(P, kMaskAll, D.psi_in, D.psi_out, 1, 1, ip, j, k, d, cx, cy, cz);
        const long long o = ((long long)ip * Ny + j) * Nz + k;
        const float imx = MC ? D.im[o] : 1.0f;
        const float imy = MC == 3 ? D.im[N + o] : imx;
        const float imz = MC == 3 ? D.im[2 * N + o] : imx;
        float hx = hp[m][0] - (c * cx) * imx;
        float hy = hp[m][1] - (c * cy) * imy;
        float hz = hp[m][2] - (c * cz) * imz;
        apply_sources<false>(P, kMaskAll, D.ie, D.im, 1, D.n, false, ip, j, k, o, 1, MC, hx, hy, hz);
        D.H_out[o] = hx;
        D.H_out[N + o] = hy;
        D.H_out[2 * N + o] = hz;
        if (D.snap && in_box(ip, j, k, D.slo, D.shi)) {
          const long long sy = D.shi[1] - D.slo[1], sz = D.shi[2] - D.slo[2];
          const long long sc = (long long)(D.shi[0] - D.slo[0]) * sy * sz;
          const long long sq = ((long long)(ip - D.slo[0]) * sy + (j - D.slo[1])) * sz + (k - D.slo[2]);
          D.snap[sq] = hp[m][0];
          D.snap[sc + sq] = hp[m][1];
          D.snap[2 * sc + sq] = hp[m][2];
        }
        if (D.record) {
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
            re[rq] = D.E_out[o];
            re[pc + rq] = D.E_out[N + o];
            re[2 * pc + rq] = D.E_out[2 * N + o];
            rh[rq] = hx;
            rh[pc + rq] = hy;
            rh[2 * pc + rq] = hz;
          }
        }
      }
      if (po[m] && i < x1) {
        const int je = (j - j0) * Nz + k;  // row of (j, k) in Es
        const float ex = Es[je], ey = Es[strideE + je], ez = Es[2 * strideE + je];
        const float ex_jp = Es[je + Nz], ez_jp = Es[2 * strideE + je + Nz];
        const float ex_kp = k + 1 < Nz ? Es[je + 1] : 0.f;
        const float ey_kp = k + 1 < Nz ? Es[strideE + je + 1] : 0.f;
        dv[m][0] = ez_jp - ez;  // dyEz
        dv[m][1] = ey_kp - ey;  // dzEy
        dv[m][2] = ex_kp - ex;  // dzEx
        dv[m][3] = ex_jp - ex;  // dyEx
        ep[m][0] = ey;
        ep[m][1] = ez;
      }
      if (pv[m]) {
        const int jh = (j - j0 + 1) * Nz + k;
        hp[m][0] = Hs[jh];
        hp[m][1] = Hs[strideH + jh];
        hp[m][2] = Hs[2 * strideH + jh];
      }
    }
    __syncthreads();
  }
}

// Tile shape for a grid: TY rows so that (TY+1)*nz positions fit M per thread; the LDS footprint
// 12*nz*(2*TY+3) bytes must fit 64 KiB. Returns false when the grid does not suit the fused path.
bool fused_config(const StepStatic& P, int target_blocks, FusedDyn& D, size_t& lds_bytes) {
  constexpr int M = kFusedM;
  int TY = (kFusedThreads * M) / P.nz - 1;
  if (TY < 1) return false;
  if (TY > P.ny) TY = P.ny;
  while (TY >= 1 && 12ll * P.nz * (2 * TY + 3) > 65536) --TY;
  if (TY < 1) return false;
  D.TY = TY;
  D.ntiles = (P.ny + TY - 1) / TY;
  int nchunks = (target_blocks + D.ntiles - 1) / D.ntiles;
  if (nchunks < 1) nchunks = 1;
  if (nchunks > P.nx) nchunks = P.nx;
  D.chunk = (P.nx + nchunks - 1) / nchunks;
  D.nchunks = (P.nx + D.chunk - 1) / D.chunk;
  D.nblocks = (long long)D.ntiles * D.nchunks;
  D.nblocks_pad = pad_blocks(D.nblocks);
  lds_bytes = 12ull * P.nz * (2 * TY + 3);
  return true;
}

hipError_t launch_fused(const StepStatic& P, const FusedDyn& D, int ec, int mc, size_t lds, hipStream_t s) {
  const dim3 g((unsigned)D.nblocks_pad), b(kFusedThreads);
  constexpr int M = kFusedM;
  if (ec == 3) {
    if (mc == 3) hipLaunchKernelGGL((k_fused<3, 3, M>), g, b, lds, s, P, D);
    else if (mc == 1) hipLaunchKernelGGL((k_fused<3, 1, M>), g, b, lds, s, P, D);
    else hipLaunchKernelGGL((k_fused<3, 0, M>), g, b, lds, s, P, D);
  } else {
    if (mc == 3) hipLaunchKernelGGL((k_fused<1, 3, M>), g, b, lds, s, P, D);
    else if (mc == 1) hipLaunchKernelGGL((k_fused<1, 1, M>), g, b, lds, s, P, D);
    else hipLaunchKernelGGL((k_fused<1, 0, M>), g, b, lds, s, P, D);
  }
  return hipGetLastError();
}

}  // namespace fdtdx
