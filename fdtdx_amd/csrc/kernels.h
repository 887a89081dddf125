// Launchers shared between the kernel translation units and the C-ABI in plan.hip.
#pragma once
#include "engine_common.h"

namespace fdtdx {

struct DetK {
  int kind;
  int lo[3], hi[3];
  int reduce;
  int accum;   // add the reduced values to the state (the detector before it stored first)
  int inv;     // recorded by the time-reversed run (fdtdx_reverse), not by the forward
  int exact;
  int ncomp;
  int comps[6];
  int axis;
  float sign;
  float weight;
  int nfreq;
  float scale;
  long long cells;
  const float* cos_t;  // device (nfreq, T)
  const float* sin_t;
  const float* window; // device T
  const int* tti;      // device time_to_idx (T): record slot of step n, or -1
  int cplx;            // complex fields: a reduced phasor sums the real, then the imaginary parts
};

struct DetDyn {
  const float* E;
  const float* H;
  const float* Hprev;  // snapshot over [slo, shi)
  int slo[3], shi[3];
  const float* ie;
  const float* im;
  int ec, mc;
  void* state;         // detector state base
  int slot;            // record slot (time_to_idx[n])
  int n;
  double* partial;     // per-block partial sums (reductions): value v of block b at v * kDetMaxBlocks + b
  // complex-field plans: the imaginary parts of E, H and the snapshot (null for real plans)
  const float* Ei;
  const float* Hi;
  const float* Hprevi;
};

constexpr int kDetMaxBlocks = 1024;  // blocks of one detector launch (detector_blocks)
constexpr int kDetGroup = 8;         // detectors of one grouped launch (the kernel arguments stay below 4 KB)

// The detectors updated at one step, in one launch: detector a of the group is entry q[a] of the
// plan's device DetK array, records into slot[a] and owns blocks [start[a], start[a + 1]) (its
// single launch's block count, so partial sums and reduced values are unchanged). The grouped
// finalize runs one block per chain head fin[b] (a reducing detector that does not accumulate),
// which then also finalizes the accumulating detectors that follow it.
struct DetGroup {
  int count, nfin;
  int q[kDetGroup];
  int slot[kDetGroup];
  int start[kDetGroup + 1];
  int fin[kDetGroup];
  DetDyn dd[kDetGroup];
};
constexpr int kDetMaxValues = 6;     // reduced values per step: 1 (energy), 1 or 3 (flux), one per component

// Reduced values a detector sums over its box per step, 0 when it keeps per-cell values.
__host__ __device__ inline int reduced_values(const DetK& K) {
  if (!K.reduce) return 0;
  if (K.kind == FDTDX_DET_ENERGY) return K.reduce == 1 ? 1 : 0;  // 2: means along an axis, no partials
  if (K.kind == FDTDX_DET_FLUX) return K.ncomp == 3 ? 3 : 1;
  if (K.kind == FDTDX_DET_PHASOR && K.cplx) return 2 * K.ncomp;
  return K.ncomp;
}

// ---------------------------------------------------------------- multi-instance batch
// A batch advances many independent solves per launch. Entry e of the device table holds one
// instance's plan tables and buffers; the persistent sweep kernels walk the concatenated tile
// space [tile_start[e], tile_start[e+1]) of all entries. Step n of entry e in a launch is
// base[slot] + r, where base is the per-slot device step counter and r the step inside the
// current chunk; an entry whose n has reached `stop` is skipped.
constexpr int kMaxBatch = 64;

struct BatchInst {
  StepStatic P;
  StepDyn D;             // snap is set whenever the instance has detectors
  char* rec_base;        // interface record slot 0 (null: off); step n at slot P.rec_slot[n]
  int stop;              // first step index not to run
  int slot;              // index into the step counter array
  int ec, mc;
  int ndet;
  int det_nb[FDTDX_MAX_DETECTORS];  // blocks per detector (detector_blocks)
  DetK det[FDTDX_MAX_DETECTORS];
  DetDyn dd[FDTDX_MAX_DETECTORS];   // per-detector buffers (n and slot are per launch)
};

struct BatchTable {
  int ninst;
  float blowup;                         // |reduced detector value| above this flags the slot (0: off)
  int* flags;                           // per-slot health flags (device)
  long long tile_start[kMaxBatch + 1];  // sweep tiles, prefix over entries
  int det_start[kMaxBatch + 1];         // detector blocks, prefix over entries
  int fin_start[kMaxBatch + 1];         // reducing detectors, prefix over entries
  BatchInst inst[kMaxBatch];
};

// Health flag bits of a batch slot, raised by the detector finalisation.
constexpr int kFlagNonFinite = 1;
constexpr int kFlagBlowup = 2;

struct BatchSet {               // counters to (re)initialise (and flags to clear) before a chunk
  int count;
  int slot[kMaxBatch];
  int value[kMaxBatch];
};

int batch_sweep_grid(long long tiles);  // blocks of a batch sweep over `tiles` tiles
hipError_t launch_batch_set(int* base, int* flags, const BatchSet& set, hipStream_t s);
hipError_t launch_batch_sweeps(const BatchTable* tab, const int* base, int r, int ec, int mc, int grid,
                               hipStream_t s);
hipError_t launch_batch_detectors(const BatchTable* tab, const int* base, int r, int grid, bool finalize,
                                  hipStream_t s);
hipError_t launch_batch_advance(const BatchTable* tab, int* base, int steps, hipStream_t s);

// Fused single-pass step (fused_kernels.hip): kFusedThreads threads, kFusedM positions per thread
// per plane.
constexpr int kFusedThreads = 1024;
constexpr int kFusedM = 2;

struct FusedDyn {
  const float* E_in;
  const float* H_in;
  float* E_out;
  float* H_out;
  const float* ie;
  const float* im;
  const float* psi_in;
  float* psi_out;
  float* ade;
  char* record;
  float* snap;
  int slo[3], shi[3];
  int n;
  int TY, chunk, ntiles, nchunks;
  long long nblocks, nblocks_pad;
};

// Register-march single-pass step (march_kernels.hip): kMarchThreads threads = independent
// wavefronts, each owning W z-lanes x TY y-rows x XC x-planes.
constexpr int kMarchThreads = 256;

struct MarchDyn {
  const float* E_in;
  const float* H_in;
  float* E_out;
  float* H_out;
  const float* ie;
  const float* im;
  const float* psi_in;
  float* psi_out;
  float* ade;
  char* record;
  float* snap;
  int slo[3], shi[3];
  int n;
  int W, nzc, TY, nyt, XC, nxc;
  long long ntiles, nblocks, nblocks_pad;
};

bool march_config(const StepStatic& P, int ty, int xc, MarchDyn& D);
hipError_t launch_march(const StepStatic& P, const MarchDyn& D, int ec, int mc, hipStream_t s);

bool fused_config(const StepStatic& P, int target_blocks, FusedDyn& D, size_t& lds_bytes);
hipError_t launch_fused(const StepStatic& P, const FusedDyn& D, int ec, int mc, size_t lds, hipStream_t s);

hipError_t launch_update_E(const StepStatic& P, const StepDyn& D, int ec, hipStream_t s);
hipError_t launch_update_H(const StepStatic& P, const StepDyn& D, int mc, hipStream_t s);
// Full-tensor plans with conductivity: the per-cell update matrices of the conductivity boxes
// (D.ab) from the current inverse material arrays; once per call, before the first half-step.
hipError_t launch_tensor_ab(const StepStatic& P, const StepDyn& D, hipStream_t s);
hipError_t launch_update_E4(const StepStatic& P, const StepDyn& D, int ec, hipStream_t s);
hipError_t launch_update_H4(const StepStatic& P, const StepDyn& D, int mc, hipStream_t s);
hipError_t launch_restore(const StepStatic& P, const StepDyn& D, long long iface_cells, hipStream_t s);
hipError_t launch_record(const StepStatic& P, const StepDyn& D, long long iface_cells, hipStream_t s);
hipError_t launch_reverse(const StepStatic& P, const StepDyn& D, int ec, int mc, hipStream_t s);
hipError_t launch_detector(const StepStatic& P, const DetK& K, const DetDyn& D, int nblocks, hipStream_t s);
hipError_t launch_detector_group(const StepStatic& P, const DetK* Kdev, const DetGroup& G, hipStream_t s);
int detector_blocks(const DetK& K);

struct AdjDyn {
  float* lamE;
  float* lamH;
  const float* lpsi_in;  // lambda psi read (step n+1)
  float* lpsi_out;       // lambda psi written (step n)
  float* grad;
  const float* H;        // reconstructed H^n
  const float* E;        // reconstructed E^n (read in conducting cells)
  const float* psi;      // frozen psi
  const float* ie;
  const float* im;
  float* W;              // full-tensor plans: transposed-tensor cotangent on a curl, (3, N)
  float* K;              // full-tensor permittivity: K^{n+1/2} of every cell, (3, N)
  float* Erev;           // fused reverse (non-null): E^{n+1} is reversed to E^n in place by pass A
  int n;                 // the step being undone (sources of the fused reverse)
};

hipError_t launch_adjoint(const StepStatic& P, const AdjDyn& A, int ec, int mc, hipStream_t s);
// The reverse H half-step only (the E half-step is then fused into the adjoint's pass A).
hipError_t launch_reverse_H(const StepStatic& P, const StepDyn& D, int mc, hipStream_t s);

// Detector cotangent injection for reverse step n.
struct DetAdjDyn {
  const float* E1;     // E^{n+1} snapshot over the snapshot box
  const float* H1;     // H^{n+1} snapshot
  const float* H0;     // H^n (full field)
  int slo[3], shi[3];
  const float* ie;
  const float* im;     // inverse permeability (mc > 0)
  int ec, mc;
  float* grad;         // dL/d inv_eps (energy detectors read inv_eps directly)
  const void* cot;     // detector cotangent base
  int slot;
  int n;
  float* g;            // workspace: 6 * region cells (interpolated-space cotangents)
  float* half_vH;      // workspace: 3 * expanded-region cells, 1/2 of the H cotangent
  float* lamE;
  float* lamH;
  int elo[3], ehi[3];  // expanded region
};

hipError_t launch_detector_adjoint(const StepStatic& P, const DetK& K, const DetAdjDyn& D, hipStream_t s);
hipError_t launch_add_half_vH(const StepStatic& P, const DetAdjDyn& D, hipStream_t s);
hipError_t launch_zero_box(const StepStatic& P, float* F, const int* lo, const int* hi, hipStream_t s);
// Snapshot of src over box [lo, hi) into dst; with src1 also that of src1 into dst1, in the same launch.
hipError_t launch_snapshot(const StepStatic& P, const float* src, float* dst, const int* lo, const int* hi,
                           hipStream_t s, const float* src1 = nullptr, float* dst1 = nullptr);

}  // namespace fdtdx
