// Shared device-side structures of the FDTD engine (gfx950 / CDNA4).
//
// Layout in HBM: fields are (3, nx, ny, nz) float32 with z contiguous, so a 64-lane wavefront
// reads 256 contiguous bytes of one component. CPML psi state is packed per slab
// [psiE1, psiE2, psiH1, psiH2] each over the slab box; ADE polarisation is packed per box
// [slot 0|1|2][pole][comp][box cells]; slot n%3 holds P^n, so a step reads two slots and writes
// the third.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fdtdx_hip.h"

namespace fdtdx {

constexpr int kBlock = 256;  // threads per workgroup = 4 wavefronts of 64
constexpr int kXcds = 8;     // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kXcdChunk = 4; // consecutive tiles per XCD turn of the block order (logical_block)

// A grid of n tiles rounded up to whole turns of the block order.
inline long long pad_blocks(long long n) {
  const long long q = (long long)kXcds * kXcdChunk;
  return (n + q - 1) / q * q;
}

struct SlabK {
  int axis, lo, hi, L;
  int sx, sy, sz;          // slab box shape
  int kappa_one;
  long long cells;         // sx*sy*sz
  long long psi_off;       // float offset of psiE1 in the packed psi buffer
  long long rec_off_E;     // byte offset of this slab's E interface layer inside one record slot
  long long rec_off_H;
  int fmt_E, fmt_H;        // record formats of the two layers (FDTDX_REC_*)
  int iface;               // index of the interface layer along axis
  int pn0, pn1;            // interface plane dims (the two other axes, ascending)
  const float* coef;       // 6*L: bE aE ikm1E bH aH ikm1H
};

struct SrcK {
  int kind, field, comp;
  int lo[3], hi[3];
  int target[2], src[2];
  float sign[2];
  int py, pz;              // plane box dims along y, z (for profile indexing)
  long long pcells;
  const float* table;      // POINT: T; PLANE: 3*T
  const float* profile;    // PLANE: 3*pcells
};

struct AdeK {
  int lo[3], hi[3];
  int poles, ccomp;
  int c3comp;              // components of c3: ccomp, or 9 (coupling tensors; full-tensor plans)
  int by, bz;
  long long cells;
  long long state_off;     // float offset in ade_state
  const float* coef;       // c1, c2 [poles][ccomp][cells], then c3 [poles][c3comp][cells]
};

struct WallK {
  int kind, axis, index;
};

struct LossK {             // conductivity box; zero conductivity outside it
  int comps;               // 0 (none), 1, 3 or 9 (a full tensor, row-major; full-tensor plans only)
  int lo[3], hi[3];
  int by, bz;
  long long cells;
  const float* sig;        // [comps][cells], scaled by c0*dt/courant
};

struct StepStatic {
  int nx, ny, nz;
  int T;
  long long plane;         // ny*nz
  long long N;             // nx*ny*nz
  float c;
  int bpp;                 // blocks per x-plane
  long long nblocks;       // nx*bpp
  long long nblocks_pad;   // pad_blocks(nblocks)
  int nslab;
  SlabK slab[FDTDX_MAX_CPML];
  int nsrc;
  SrcK src[FDTDX_MAX_SOURCES];
  int nade;
  AdeK ade[FDTDX_MAX_ADE];
  int nwall;
  WallK wall[FDTDX_MAX_WALLS];
  int wrap[3];             // periodic axes
  // Complex-field plans with a Bloch phase (bloch = 1): the neighbour read across the min face of
  // a periodic axis a is the far face's value times exp(-i theta_a,min), across the max face the
  // near face's value times exp(+i theta_a,max) (bloch.py:61-97); bcos/bsin[a][side] hold
  // cos and sin of theta (1 and 0 on axes without a phase).
  int bloch;
  float bcos[3][2], bsin[3][2];
  int mirror[3];           // electric symmetry plane at the min face (detector stencils only)
  LossK loss[2];           // 0: electric conductivity, 1: magnetic conductivity
  // tsig[w] = 1: the full-tensor half-step of field w (0: E, 1: H) has conductivity, so inside its
  // box F' = A avg(F) +- B avg(K) with per-cell matrices (update.py:356-492, 752-822)
  int tsig[2];
  float eta0;              // free-space impedance mu0*c0 rounded to float32
  long long rec_per_step;  // interface values per record step
  long long rec_slot_bytes;  // bytes of one record slot
  const int* rec_slot;     // T entries: slot of step n, -1 when not kept (null: slot n for step n)
  int rec_filter_first;    // interpolate in float32 (1) or in the layer's format (0)
  int rec_narrow;          // some record layer is stored in a narrow format
  // Per logical tile (kBlock consecutive cells of one x-plane): which descriptors can touch it.
  // Bits [0, 6) CPML slabs (footprint widened by one cell, for the adjoint's neighbour terms),
  // [6, 14) sources, [14, 18) ADE boxes, bit 18 the H_prev snapshot box, [19, 25) PEC / PMC faces
  // (widened by one cell), 25 and 26 the electric and magnetic conductivity boxes (widened by one
  // cell, across periodic faces too). A superset: the per-cell tests still decide. Null means
  // every bit set.
  const uint32_t* tile_mask;
  // Forward sweeps visit each x-plane in a permuted order, in groups of zg rows: first the
  // group's cells with zlo <= k < zhi (row by row), then its cells of the z-normal CPML slabs
  // (k < zlo or k >= zhi, row by row), so a wavefront runs the CPML code for all or none of its
  // lanes while both parts of a row are written close together in time. fwd_mask holds the tile
  // masks of that order. zlo = 0, zhi = nz is the plain row-major order.
  int zlo, zhi, zg;
  const uint32_t* fwd_mask;
  // Quad sweeps (nz a multiple of 4): each lane updates 4 consecutive z cells with dwordx4 loads
  // and stores; a quad tile is kBlock quads (4 kBlock cells) of one x-plane, visited in the same
  // row-grouped order as above at quad granularity (vzlo, vzhi in quads; plain row-major when the
  // z-slab edges are not multiples of 4). vq = 0: the one-cell sweeps run.
  int vq;
  int vzlo, vzhi, vzg;
  int vbpp;                // quad tiles per x-plane
  long long vnblocks, vnblocks_pad;
  const uint32_t* vec_mask;  // per quad tile, as fwd_mask
  long long vnslow;          // quad tiles whose E half-step runs one cell per thread (vlistE bits) ...
  const int* vslow;          // ... and their indices
  long long vnslowH;         // the same for the H half-step
  const int* vslowH;
  uint32_t vlistE, vlistH;   // the mask bits that list a quad tile for the E / H half-step
  // Per-axis CPML (acax = 1): every slab spans its face, at most one slab per side of an axis and
  // no two overlapping, and the x slabs before the y slabs before the z slabs in object order (the
  // order in which the corrections are added); the one-cell reverse and adjoint sweeps then find
  // a cell's slab per axis from its coordinate. The quad sweeps (cax = 1) also need no quad
  // meeting both z slabs. axs[a][side] is the slab index (-1: none); axord lists the axes in the
  // order of their slab indices.
  int cax, acax;
  int xorder;                // block order (logical_block): 0 per-XCD ranges, 1 XCD turns
  int axs[3][2];
  int axord[3];
  struct AxSlab {            // the slab of side 0 (min) / 1 (max) of an axis, as SlabK
    int lo, hi, L, kappa_one;
    int aligned;             // z slabs: lo and L multiples of 4 (whole quads move psi as dwordx4)
    long long cells, psi_off;
    const float* coef;
  } axd[3][2];
};

// Row j and column k of position p of a plane of ny rows of nz positions, in groups of zg rows:
// the group's positions zlo <= k < zhi row by row, then its other positions row by row.
__host__ __device__ __forceinline__ void perm_pos(int p, int ny, int nz, int zlo, int zhi, int zg, int& j, int& k) {
  const int W = zhi - zlo;
  const int gs = zg * nz;               // positions per row group
  const int g = p / gs;
  const int r = p - g * gs;
  const int j0 = g * zg;
  const int rows = ny - j0 < zg ? ny - j0 : zg;
  if (r < rows * W) {
    j = r / W;
    k = zlo + (r - j * W);
  } else {  // the z-slab positions of the group's rows
    const int r2 = r - rows * W, Wz = nz - W;
    j = r2 / Wz;
    const int kk = r2 - j * Wz;
    k = kk < zlo ? kk : kk + W;
  }
  j += j0;
}

// Row j and column k of position p of an x-plane in the forward sweeps' order.
__host__ __device__ __forceinline__ void perm_cell(const StepStatic& P, int p, int& j, int& k) {
  perm_pos(p, P.ny, P.nz, P.zlo, P.zhi, P.zg, j, k);
}

// Row j and first column k of quad p of an x-plane in the quad sweeps' order.
__host__ __device__ __forceinline__ void perm_quad(const StepStatic& P, int p, int& j, int& k) {
  int kq;
  perm_pos(p, P.ny, P.nz / 4, P.vzlo, P.vzhi, P.vzg, j, kq);
  k = 4 * kq;
}

constexpr int kMaskSrc = 6;
constexpr int kMaskAde = 14;
constexpr int kMaskSnap = 18;
constexpr int kMaskWall = 19;  // bits [19, 25): PEC / PMC faces (footprint widened by one cell)
constexpr int kMaskLoss = 25;  // bits 25, 26: conductivity boxes (footprint widened by one cell)
constexpr uint32_t kMaskAll = 0xffffffffu;

struct StepDyn {
  float* E;
  float* H;
  const float* ie;
  const float* im;
  float* psi;
  float* ade;
  char* record;            // slot of step n's record (nullptr: none); the reverse pass's earlier slot
  const char* rec_b;       // reverse pass, unkept step: the later slot
  float rec_f;             // reverse pass, unkept step: interpolation factor
  int rec_interp;          // reverse pass: 1 when step n is interpolated from record and rec_b
  float* snap;             // H_prev snapshot (nullptr: none)
  float* K;                // full-tensor plans: the curl (with CPML) of the current half-step, (3, N)
  int slo[3], shi[3];      // snapshot box
  int n;
  // Full-tensor plans with conductivity: the field of the current half-step as it was before the
  // update (the A avg(F) term reads its neighbours), and per conductivity-box cell the matrices
  // A, B, A_rev, B_rev (36 floats, row-major, [36][cells]) of E (ab[0]) and H (ab[1]).
  float* Fo;
  const float* ab[2];
  // Complex-field plans store the real part of every field array, then the imaginary part; a
  // sweep updates one part: cpart = +1 the real part (the imaginary part lies 3 N floats after the
  // pointers above), -1 the imaginary part (the real part 3 N floats before), 0 a real plan.
  int cpart;
};

// The value a complex-field sweep reads across the min (side 0) or max (side 1) face of periodic
// axis a: v (this part, at the wrapped cell) rotated by the face's Bloch phase, w the other part
// at the same cell. Re(z e^{-+i t}) = cos t Re z +- sin t Im z, Im(z e^{-+i t}) = cos t Im z -+ sin t Re z.
__device__ __forceinline__ float bloch_ghost(const StepStatic& P, int a, int side, float v, float w, int cpart) {
  const float sw = (float)cpart * P.bsin[a][side];
  return side == 0 ? P.bcos[a][0] * v + sw * w : P.bcos[a][1] * v - sw * w;
}

// Loads and stores of the arrays a sweep touches once per cell (the updated field, inv_eps /
// inv_mu, psi). Plain cached accesses: with the nontemporal hint the 400x400x200 forward measured
// 0.381 against 0.331 ms per E sweep (and 0.358 against 0.297 ms per H sweep) on one box.
__device__ __forceinline__ float ldg(const float* p) { return *p; }
__device__ __forceinline__ void stg(float* p, float v) { *p = v; }

// Block order: hardware block h lands on XCD h % 8. Order 1: the XCDs take turns over runs of
// kXcdChunk consecutive tiles (the r-th block of XCD x takes tile ((r / kXcdChunk) * 8 + x) *
// kXcdChunk + r % kXcdChunk), so the whole device sweeps one front of the grid and the XCDs share
// its HBM pages, re-reading neighbour planes through the last-level cache. Order 0: each XCD sweeps
// a contiguous range of tiles, whose neighbour planes stay in its own L2. The plan takes order 1
// when a sweep's arrays exceed the 256 MiB last-level cache (StepStatic::xorder). Measured: the
// 400x400x200 forward (1.4 GB per E sweep) 0.537 ms per step in order 1 against 0.581 ms in
// order 0; the 200x200x100 forward (160 MB) 0.099 against 0.094 ms. Grids are padded to whole
// turns (pad_blocks, also a multiple of kXcds); blocks past the last tile return.
__device__ __forceinline__ long long logical_block(int order, long long nblocks_pad) {
  const long long h = blockIdx.x;
  const long long x = h % kXcds, r = h / kXcds;
  if (order == 0) return x * (nblocks_pad / kXcds) + r;
  return ((r / kXcdChunk) * kXcds + x) * kXcdChunk + r % kXcdChunk;
}

// Offset of the backward (prev_off) or forward (next_off) neighbour of a cell at `coord` along
// axis a: one stride, or across the domain on a periodic axis; false beyond a non-periodic face
// (zero halo).
// X = false compiles the wrap out (plans without periodic axes).
template <bool X = true>
__device__ __forceinline__ bool prev_off(const StepStatic& P, int a, int coord, long long& off) {
  const long long stride = a == 0 ? P.plane : (a == 1 ? (long long)P.nz : 1ll);
  const int n = a == 0 ? P.nx : (a == 1 ? P.ny : P.nz);
  if (coord > 0) { off = -stride; return true; }
  off = (long long)(n - 1) * stride;
  return X && P.wrap[a] != 0;
}

template <bool X = true>
__device__ __forceinline__ bool next_off(const StepStatic& P, int a, int coord, long long& off) {
  const long long stride = a == 0 ? P.plane : (a == 1 ? (long long)P.nz : 1ll);
  const int n = a == 0 ? P.nx : (a == 1 ? P.ny : P.nz);
  if (coord < n - 1) { off = stride; return true; }
  off = -(long long)(n - 1) * stride;
  return X && P.wrap[a] != 0;
}

// Coordinate of the neighbour at coord + d (d = -1 or +1) along an axis of n cells, wrapped.
__device__ __forceinline__ int wrap_coord(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

__device__ __forceinline__ bool in_box(int i, int j, int k, const int* lo, const int* hi) {
  return i >= lo[0] && i < hi[0] && j >= lo[1] && j < hi[1] && k >= lo[2] && k < hi[2];
}

}  // namespace fdtdx
