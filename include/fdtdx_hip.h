/*
 * fdtdx_hip.h -- C ABI of the MI355X FDTD time-loop engine (libfdtdx_hip.so).
 *
 * The reference (ymahlau/fdtdx 0.6.2) has no native boundary: its time loop is the body of an
 * equinox while-loop traced into one XLA executable. The entry points below replace that loop
 * body and are what a JAX FFI custom call, a ctypes stub or a torch extension binds:
 *
 *   fdtdx_forward          replaces the while-loop over forward() in checkpointed_fdtd
 *                          (src/fdtdx/fdtd/fdtd.py:482-493), custom_fdtd_forward (fdtd.py:574-581)
 *                          and the segmented forward of reversible_fdtd (fdtd.py:143-166); one
 *                          iteration is forward() (src/fdtdx/fdtd/forward.py:83-156).
 *   fdtdx_reverse_adjoint  replaces the reverse while-loop of reversible_fdtd's fdtd_bwd
 *                          (fdtd.py:215-251, 318-323): backward() (fdtd/backward.py:62-135) plus
 *                          jax.vjp of forward() at the reconstructed state.
 *   fdtdx_plan_create      replaces the host-side lowering that place_objects/apply_params leave
 *                          in the ObjectContainer (fdtd/initialization.py:100-521): CPML profiles
 *                          (objects/boundaries/perfectly_matched_layer.py:97-136), source tables
 *                          (objects/sources/dipole.py:150-234, tfsf.py:193-409), detector schedules
 *                          (objects/detectors/detector.py:195-244).
 *
 * Conventions: every function returns an int status (0 = FDTDX_OK); fdtdx_last_error() returns
 * a thread-local message for the last failure on the calling thread. Arrays are C-order with the
 * last axis (z) contiguous: fields are (3, nx, ny, nz) float32. Host pointers in descriptors are
 * read during fdtdx_plan_create only. Device buffers are owned by the caller; the plan owns only
 * its immutable device tables. A plan may be shared by threads; calls with disjoint buffers on
 * different streams may run concurrently.
 */
#ifndef FDTDX_HIP_H
#define FDTDX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FDTDX_OK = 0,
  FDTDX_EINVAL = 1,  /* bad descriptor or buffer (message says which) */
  FDTDX_EHIP = 2,    /* a HIP runtime call failed */
  FDTDX_ENOTSUP = 3, /* configuration outside what the engine implements */
};

#define FDTDX_MAX_CPML 6
#define FDTDX_MAX_SOURCES 8
#define FDTDX_MAX_ADE 4
#define FDTDX_MAX_POLES 4
#define FDTDX_MAX_DETECTORS 16
#define FDTDX_MAX_WALLS 6

/* One CPML object (PerfectlyMatchedLayer, perfectly_matched_layer.py:12-190). The slab spans
 * [lo, hi) along `axis` and the full domain along the other two axes
 * (objects/boundaries/initialization.py:598-619). Coefficient vectors have hi-lo entries in grid
 * order; ikm1 = 1/kappa - 1 (only read when kappa_is_one == 0). */
typedef struct fdtdx_cpml_desc {
  int32_t axis;
  int32_t side; /* 0 = min ("-"), 1 = max ("+") */
  int32_t lo, hi;
  int32_t kappa_is_one;
  const float *b_E, *a_E, *ikm1_E;
  const float *b_H, *a_H, *ikm1_H;
} fdtdx_cpml_desc;

/* A PEC or PMC face (PerfectElectricConductor, pec.py:70-77; PerfectMagneticConductor,
 * pmc.py:62-75): after every E (PEC) or H (PMC) update, in the forward and in the reverse
 * reconstruction (update.py:521, 683, 851, 1004), the two field components tangential to `axis`
 * are set to zero on the one-cell layer at `index` along `axis`. */
enum { FDTDX_WALL_PEC = 0, FDTDX_WALL_PMC = 1 };
typedef struct fdtdx_wall_desc {
  int32_t kind;
  int32_t axis;
  int32_t index;
} fdtdx_wall_desc;

enum { FDTDX_SRC_POINT = 0, FDTDX_SRC_PLANE = 1 };

/* A source hook (Source.update_E/update_H, objects/sources/source.py:169-213), applied after the
 * curl update of its field in list order.
 *  POINT: field[component][cell] += table[n]      (PointDipoleSource.update_E, dipole.py:195-234;
 *         the host folds -c*amplitude*A(t)*inv_eps_local into the table)
 *  PLANE: for t in {0,1}: field[target[t]][p] += sign[t] * ((profile[src_comp[t]][p] *
 *         amp[src_comp[t]*T + n]) * courant) * mult[t](p)   over the plane box p, where mult is
 *         inv_eps[target] for E-side terms and inv_mu (1.0 when absent) for H-side terms
 *         (_tfsf_inject_E_face / _H_face, tfsf.py:193-409). */
typedef struct fdtdx_source_desc {
  int32_t kind;
  int32_t field; /* 0 = E, 1 = H */
  int32_t component;
  int32_t lo[3], hi[3]; /* POINT: the cell is lo; PLANE: the face box */
  int32_t target[2], src_comp[2];
  float sign[2];
  const float* table;   /* POINT: T floats; PLANE: amp, 3*T floats */
  const float* profile; /* PLANE: 3 * (plane cells) floats, C-order over the box */
} fdtdx_source_desc;

/* A box of Lorentz/Drude ADE cells (update.py:316-347). Coefficients c1, c2 are
 * (num_poles, coef_comps, box cells) and c3 is (num_poles, coupling_comps, box cells); coef_comps
 * is 1 (isotropic) or 3 (per axis); coupling_comps is 0 (= coef_comps) or 9 (row-major 3x3
 * coupling tensors of oriented poles, compute_pole_coefficients_tensor, dispersion.py:888-952;
 * full-tensor permittivity plans only, where the off-diagonal entries couple the neighbouring E
 * components through the symmetrized four-point means of update.py:376-414). The state of a box
 * is [3 slots][num_poles][3][box cells]; slot n%3 holds P^n at the start of step n. */
typedef struct fdtdx_ade_desc {
  int32_t lo[3], hi[3];
  int32_t num_poles;
  int32_t coef_comps;
  const float *c1, *c2, *c3;
  int32_t coupling_comps;
} fdtdx_ade_desc;

enum { FDTDX_DET_ENERGY = 0, FDTDX_DET_FIELD = 1, FDTDX_DET_FLUX = 2, FDTDX_DET_PHASOR = 3 };

/* A detector (objects/detectors/) over box [lo, hi). State layouts (caller allocated):
 *  ENERGY (energy.py:90-143): reduce_volume 1: (n_rec, 1); 0: (n_rec, box); 2: (n_rec, plane) float32,
 *         the mean along `axis` (as_slices with the collapsed-axis mean), plane the two other box
 *         axes in increasing order
 *  FIELD  (field.py:31-63):   reduce_volume ? (n_rec, num_components) : (n_rec, num_components, box) float32
 *  FLUX   (poynting_flux.py:171-195): reduce_volume ? (n_rec, m) : (n_rec, m, box) float32, m = 3 with
 *         num_components == 3 (keep_all_components: the x, y, z components), else m = 1 (the
 *         propagation axis)
 *  PHASOR (phasor.py:190-235): (1, num_freqs, num_components, reduce_volume ? () : box) complex64
 *         (re, im pairs)
 * reduce_volume sums energy and flux over the box with the weight, and averages field and phasor
 * values over it (the volume-weighted mean of detector.py:95-109 on a uniform grid).
 * time_to_idx[n] is the record slot of step n or -1 when the detector is off. */
typedef struct fdtdx_detector_desc {
  int32_t kind;
  int32_t lo[3], hi[3];
  int32_t reduce_volume;
  int32_t exact_interpolation;
  int32_t num_components;
  int32_t components[6]; /* 0..5 = Ex, Ey, Ez, Hx, Hy, Hz */
  int32_t axis;          /* FLUX propagation axis */
  float sign;            /* FLUX: +1 for direction "+", -1 for "-"; PHASOR: -1 when inverse */
  float weight;          /* cell volume (ENERGY) or face area (FLUX) */
  int32_t num_freqs;
  float scale; /* PHASOR static scale */
  const int32_t* time_to_idx; /* T entries */
  const float *cos_table, *sin_table; /* PHASOR: (num_freqs, T) */
  const float* window;                /* PHASOR: T */
  /* ENERGY / FLUX with reduce_volume: add the reduced values to the state instead of storing
   * them. The caller passes the state buffer of the detector before it, of the same kind, value
   * count and time_to_idx, which stores first; a closed surface (the net flux through the faces
   * of a box, poynting_flux.py:196-286) is one such group of face planes. */
  int32_t accumulate;
  /* 1: an inverse detector (detector.py `inverse`), recorded by fdtdx_reverse on the time-reversed
   * run (backward.py:62-135) instead of by the forward; a PHASOR subtracts (sign -1). */
  int32_t inverse;
} fdtdx_detector_desc;

/* Conductivity inside box [lo, hi) (update.py:300-354 for sigma_E, 741-750 for sigma_H), with
 * zero conductivity outside it. sigma is (comps, box) C-order float32, already scaled by
 * c0 * dt / courant as the reference stores it (initialization.py:786, 892-930); comps is 1 or 3,
 * or 0 for no conductivity. The E update becomes E = ((1 - x) E + c curl inv_eps [+ ADE]) / (1 + x)
 * with x = c * sigma * eta0 * inv_eps / 2; the H update uses x = c * sigma / eta0 * inv_mu / 2. */
typedef struct fdtdx_loss_desc {
  int32_t comps;
  int32_t lo[3], hi[3];
  const float* sigma;
} fdtdx_loss_desc;

/* Interface-record formats: the narrow float types of interfaces/modules.py DtypeConversion.
 * Conversions round to nearest, ties to even; values beyond the largest finite one become inf
 * (BF16, F16, E5M2) or NaN (the formats without inf); the FNUZ formats have no negative zero. */
enum {
  FDTDX_REC_F32 = 0, FDTDX_REC_BF16 = 1, FDTDX_REC_F16 = 2, FDTDX_REC_E4M3FN = 3,
  FDTDX_REC_E4M3FNUZ = 4, FDTDX_REC_E5M2 = 5, FDTDX_REC_E5M2FNUZ = 6
};

/* Interface-record compression (interfaces/recorder.py:32-199 with DtypeConversion,
 * modules.py:98-161, and LinearReconstructEveryK, time_filter.py:140-257). Slab s keeps its E and
 * H interface layers in format_E[s] and format_H[s]. every_k > 1 keeps only the steps start,
 * start + k, ... and T - 1 ("slots"); the reverse pass reconstructs the other steps by linear
 * interpolation between the surrounding slots, in float32 after decoding when filter_first is
 * set (the time filter precedes the casts in the module list), else in the narrow format with
 * every operation rounded to it. All zero: every step in float32 (Recorder(modules=[])). */
typedef struct fdtdx_record_desc {
  int32_t every_k; /* 0 or 1: every step */
  int32_t start;   /* start_recording_after */
  int32_t filter_first;
  int32_t format_E[FDTDX_MAX_CPML];
  int32_t format_H[FDTDX_MAX_CPML];
} fdtdx_record_desc;

typedef struct fdtdx_plan_desc {
  int32_t nx, ny, nz;
  int32_t num_steps; /* length T of every per-step table */
  float courant;     /* c = courant_factor / sqrt(3) (config.py:165-177) */
  /* inverse permittivity components: 1 (isotropic), 3 (diagonal) or 9 (full tensor, row-major
   * (3, 3), update.py:356-492). A full tensor (here or in mu_comps) takes two launches per half-step
   * and a 6 N float workspace (9 N with conductivity); it excludes plane sources and energy detectors
   * (FDTDX_ENOTSUP), gradients with conductivity or ADE poles, and multi-instance batches. Inside a
   * conductivity box of a full-tensor half-step the update is A avg(F) +- B avg(K) with the per-cell
   * matrices of fdtd/misc.py:69-129; a conductivity tensor (comps 9) needs its field's full tensor.
   * ADE poles of a full-tensor permittivity fold their polarisation current into the curl
   * (update.py:361-423). */
  int32_t eps_comps;
  int32_t mu_comps;  /* 0 (scalar 1.0, non-magnetic), 1, 3 or 9 */
  int32_t num_cpml;
  const fdtdx_cpml_desc* cpml; /* in object order */
  int32_t num_sources;
  const fdtdx_source_desc* sources; /* in object order */
  int32_t num_ade;
  const fdtdx_ade_desc* ade;
  int32_t num_detectors;
  const fdtdx_detector_desc* detectors;
  int32_t num_walls;
  const fdtdx_wall_desc* walls; /* in object order */
  /* Periodic axes (BlochBoundary with a zero Bloch vector, bloch.py): the field halo wraps to the
   * far side of the domain (update.py:92-137). No CPML slab may lie on a periodic axis, and a
   * detector must span a periodic axis fully or stay one cell away from both of its faces. */
  int32_t periodic[3];
  fdtdx_loss_desc sigma_E; /* electric conductivity (comps = 0: none) */
  fdtdx_loss_desc sigma_H; /* magnetic conductivity (comps = 0: none) */
  fdtdx_record_desc record; /* interface-record compression (all zero: exact record) */
  /* Electric mirror-symmetry planes (config.symmetry[a] == -1) at the min face of axis a, where
   * the caller also places a PEC wall: detector stencils read below that face the parity-signed
   * mirror of the interior instead of the zero halo (pad_fields_with_symmetry_mirror,
   * update.py:139-190); components sampled on the plane (tangential E, normal H) mirror row 1 with
   * sign -1, the others row 0 with sign +1. Not on a periodic axis. */
  int32_t mirror[3];
  /* Complex fields (config.use_complex_fields, fdtd/initialization.py:580-594; needed by a
   * BlochBoundary with a nonzero Bloch vector, bloch.py:32-38): E, H, psi and ade_state hold the
   * real part followed by the imaginary part (each twice the size of a real plan; the psi and ADE
   * queries report the doubled sizes). Sources are real and drive the real part. bloch_phase[a][s]
   * is the Bloch phase k_a L_a of the min (s = 0) or max (s = 1) boundary of periodic axis a
   * (get_bloch_phase, bloch.py:140-153): across the min face the far face's field enters times
   * exp(-i phase), across the max face the near face's times exp(+i phase) (apply_pad_correction,
   * bloch.py:61-97). Detectors see the complex fields: energy |E|^2 + |H|^2, flux Re(E x conj H),
   * field detectors the real part, phasors the DFT of the complex signal. Complex plans take the
   * one-cell schedule, no interface record, reverse run or batch, and no full tensor. */
  int32_t complex_fields;
  double bloch_phase[3][2];
} fdtdx_plan_desc;

typedef struct fdtdx_plan fdtdx_plan;

/* Device buffers of one solve. psi, ade_state and record are packed; their sizes come from the
 * plan queries below. record holds the interface-record slots, slot i at byte offset
 * i * fdtdx_record_slot_bytes (only read/written when non-NULL and by the reversible path); the
 * step of each slot comes from fdtdx_record_slot_of. detector_state[d] points to detector d's
 * state. */
typedef struct fdtdx_buffers {
  float* E;
  float* H;
  const float* inv_eps;
  const float* inv_mu; /* NULL when mu_comps == 0 */
  float* psi;
  float* ade_state;
  void* const* detector_state;
  void* record;
  void* workspace;
} fdtdx_buffers;

/* Adjoint state for fdtdx_reverse_adjoint. lam_* are dL/d(state) at t1 on entry and at t0 on
 * exit; grad_inv_eps is accumulated into (+=). detector_cotangent[d] is dL/d(state of detector d)
 * in the state layout above, or NULL; a complex entry z takes (dL/dRe z, dL/dIm z), the
 * convention of torch autograd (a JAX cotangent is its complex conjugate). */
typedef struct fdtdx_adj_buffers {
  float* lam_E;
  float* lam_H;
  float* lam_psi;
  float* grad_inv_eps;
  const void* const* detector_cotangent;
} fdtdx_adj_buffers;

int fdtdx_plan_create(const fdtdx_plan_desc* desc, fdtdx_plan** out);
void fdtdx_plan_destroy(fdtdx_plan* plan);

int fdtdx_psi_floats(const fdtdx_plan* plan, size_t* out);
int fdtdx_ade_floats(const fdtdx_plan* plan, size_t* out);
int fdtdx_record_floats_per_step(const fdtdx_plan* plan, size_t* out); /* interface values per step */
/* Bytes of one record slot, the number of slots of a whole run, and the slot holding step n
 * (-1 when the time filter drops it). Without compression slot n holds step n. */
int fdtdx_record_slot_bytes(const fdtdx_plan* plan, size_t* out);
int fdtdx_record_slots(const fdtdx_plan* plan, size_t* out);
int fdtdx_record_slot_of(const fdtdx_plan* plan, int32_t step, int32_t* out);
int fdtdx_workspace_bytes(const fdtdx_plan* plan, size_t* out);
/* The forward sweep schedule chosen at plan creation (no counterpart in the reference; for tests
 * and measurement): FDTDX_SCHED_ONE (E and H sweeps, one cell per lane), FDTDX_SCHED_QUAD (E and H
 * sweeps, four z cells per lane; the default when nz is a multiple of 4 and no full tensor is
 * present), FDTDX_SCHED_FUSED or FDTDX_SCHED_MARCH (single-pass, opted in by FDTDX_SWEEP). */
enum { FDTDX_SCHED_ONE = 0, FDTDX_SCHED_QUAD = 1, FDTDX_SCHED_FUSED = 2, FDTDX_SCHED_MARCH = 3 };
int fdtdx_plan_schedule(const fdtdx_plan* plan, int32_t* out);

/* Advance steps n = t0 .. t1-1 in place (forward.py:83-156 per step). When rec != 0 the PML
 * interface layers of each kept step are written, compressed, to its slot of buffers->record
 * (update.py:1140-1178, recorder.py:70-119). */
int fdtdx_forward(const fdtdx_plan* plan, const fdtdx_buffers* buffers, int32_t t0, int32_t t1,
                  int32_t rec, void* stream);

/* fdtdx_forward without allocation (no counterpart in the reference): for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Identical steps and results; buffers->workspace is ignored and the forward's
 * workspace (the H_prev snapshot of the detector boxes and the detectors' partial sums) lives in a
 * static 32 MiB device array of the library, one per GPU, held until the process ends. Plans whose
 * forward needs more (fdtdx_workspace_bytes is an upper bound), single-pass schedules and
 * full-tensor plans are rejected. Calls that overlap on one GPU share the array and race; the call
 * allocates nothing (the array's device address is looked up once per device). */
int fdtdx_forward_noalloc(const fdtdx_plan* plan, const fdtdx_buffers* buffers, int32_t t0, int32_t t1, int32_t rec,
                          void* stream);

/* Measurement variant of fdtdx_forward (no counterpart in the reference): identical work, plus
 * HIP events on `stream` around each step's field-update launches. On return sweep_ms[0] holds
 * the summed time of the E sweeps and sweep_ms[1] that of the H sweeps over the launched steps
 * (a single-pass schedule reports its one kernel in sweep_ms[0] and 0 in sweep_ms[1]). The call
 * synchronises on the last event. */
int fdtdx_forward_timed(const fdtdx_plan* plan, const fdtdx_buffers* buffers, int32_t t0, int32_t t1,
                        int32_t rec, void* stream, double* sweep_ms);

/* Reverse steps n = t1-1 down to t0: restore interfaces from record[n], reconstruct the fields
 * with frozen CPML psi (backward.py:62-135), and propagate the adjoint through the transposed
 * step, accumulating dL/d inv_eps = sum_n c * lam_E^{n+1} * curl(H^n) (fdtd.py:215-251). */
/* The time-reversed run alone (full_backward, backward.py:18-135): for n = t1-1 down to t0,
 * restore step n's interface layers from the record, undo the H and then the E half-step with
 * frozen psi, zero E and H in every CPML slab when flags & 1 (reset_fields: the PML's
 * apply_field_reset, perfectly_matched_layer.py:227-229), and record the inverse detectors at
 * step n when flags & 2 (their time-centred H averages the H before and after the undo). */
int fdtdx_reverse(const fdtdx_plan* plan, const fdtdx_buffers* buffers, int32_t t1, int32_t t0, int32_t flags,
                  void* stream);

int fdtdx_reverse_adjoint(const fdtdx_plan* plan, const fdtdx_buffers* buffers,
                          const fdtdx_adj_buffers* adj, int32_t t1, int32_t t0, void* stream);

/* ---- Multi-instance batch (serving). No counterpart in the reference, which runs one solve per
 * jitted call (fdtd/fdtd.py:421-496). A batch advances up to FDTDX_MAX_BATCH independent solves
 * of one material class (eps_comps, mu_comps) per kernel launch: persistent sweep and detector
 * kernels walk a device table of the attached instances, each at its own step index, so the
 * per-cell arithmetic and every detector reduction are those of fdtdx_forward on the same solve.
 * Attaching or detaching changes the table; the next fdtdx_batch_advance uploads it. Plans and
 * buffers must outlive their attachment. A batch is driven from one host thread. */
#define FDTDX_MAX_BATCH 64
typedef struct fdtdx_batch fdtdx_batch;

#define FDTDX_FLAG_NONFINITE 1 /* a reduced detector value was NaN or Inf */
#define FDTDX_FLAG_BLOWUP 2    /* |reduced detector value| exceeded the batch's blowup_limit */

/* blowup_limit > 0 flags a slot whose reduced (energy / flux) detector value exceeds it in
 * magnitude; 0 disables that check. Non-finite values are always flagged. */
int fdtdx_batch_create(int32_t eps_comps, int32_t mu_comps, float blowup_limit, fdtdx_batch** out);
void fdtdx_batch_destroy(fdtdx_batch* batch);
/* Attach a solve to slot [0, FDTDX_MAX_BATCH): it will run steps t0 .. t1-1 (recording the
 * interface layers when rec != 0, as fdtdx_forward). Replaces a previous attachment of the slot. */
int fdtdx_batch_attach(fdtdx_batch* batch, int32_t slot, const fdtdx_plan* plan, const fdtdx_buffers* buffers,
                       int32_t t0, int32_t t1, int32_t rec);
int fdtdx_batch_detach(fdtdx_batch* batch, int32_t slot);
/* Enqueue `steps` steps of every attached instance on `stream` (an instance stops at its t1).
 * With use_graph != 0 the chunk is one hipGraph launch, captured on first use for this step
 * count (the stream must then not be the legacy default stream). */
int fdtdx_batch_advance(fdtdx_batch* batch, int32_t steps, int32_t use_graph, void* stream);
/* Step index the slot will have reached once the enqueued chunks complete. */
int fdtdx_batch_step_index(const fdtdx_batch* batch, int32_t slot, int32_t* n);
/* Health flags (FDTDX_FLAG_*) of all FDTDX_MAX_BATCH slots as copied back after the most recent
 * chunk; valid once that chunk has completed on its stream (e.g. after an event on it). */
int fdtdx_batch_flags(const fdtdx_batch* batch, int32_t* flags);

/* Streaming probe (measurement only; no counterpart in the reference): one launch that reads
 * `reads` arrays and writes `writes` arrays of n floats each (array r of in at in + r*n, array w of
 * out at out + w*n), `vec` (1 or 4) floats per lane per access, nontemporal loads and stores when
 * nontemporal != 0. (reads, writes) is one of (1, 1), (7, 3), (6, 3), (4, 1), (1, 0); n is a
 * multiple of 4. The HBM ceiling the field sweeps are compared with. */
int fdtdx_stream_probe(int32_t reads, int32_t writes, int32_t vec, int32_t nontemporal, const float* in, float* out,
                       long long n, void* stream);

/* Stencil probe (measurement only; no counterpart in the reference): one minimal E (which = 0) or
 * H (which = 1) half-step of an isotropic, lossless, CPML-free nx x ny x nz grid with a zero halo,
 * one cell per lane in row-major order: E and H are (3, nx, ny, nz), inv_eps (nx, ny, nz),
 * N = nx ny nz < 2^31. flags bit 0 hands each XCD a contiguous range of blocks, bit 1 makes the
 * once-read arrays nontemporal. The ceiling of the sweeps' stencil without their descriptor, mask
 * and plane-order work. */
int fdtdx_stencil_probe(int32_t which, int32_t flags, float* E, float* H, const float* inv_eps, int32_t nx,
                        int32_t ny, int32_t nz, void* stream);

const char* fdtdx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FDTDX_HIP_H */
