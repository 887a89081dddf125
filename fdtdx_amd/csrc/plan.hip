// C ABI of the FDTD engine: plan creation (host tables -> immutable device tables), the forward
// time loop and the reverse/adjoint loop. See include/fdtdx_hip.h for the contract.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <map>
#include <tuple>
#include <mutex>
#include <cstdlib>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "engine_common.h"
#include "kernels.h"

using namespace fdtdx;

namespace {

thread_local std::string g_err;
// Serialises the calls that allocate, copy synchronously or free device memory (plan and batch
// creation and destruction) against hipGraph capture: a hipFree or a synchronous copy issued by
// another thread while a batch chunk is being captured would synchronise with the capturing
// stream and invalidate the capture.
std::mutex g_device_mu;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return fail(FDTDX_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

struct fdtdx_plan {
  StepStatic P;
  int ec = 1, mc = 0;
  std::vector<void*> allocs;
  size_t psi_floats = 0, ade_floats = 0;
  long long iface_cells = 0;
  int ndet = 0;
  DetK det[FDTDX_MAX_DETECTORS];
  const DetK* det_dev = nullptr;             // device copy of det (grouped detector launches)
  std::vector<std::vector<int32_t>> tti;     // per detector time_to_idx (host)
  std::vector<std::vector<float>> window;    // per detector phasor window (host)
  // interface record (recorder.py): slots, the slot of each step (-1: not kept), and for the
  // reverse pass the earlier slot and interpolation factor of each step
  long long rec_slots = 0;
  std::vector<int32_t> rec_slot_of, rec_arr;
  std::vector<float> rec_factor;
  int has_snap = 0;
  int slo[3] = {0, 0, 0}, shi[3] = {0, 0, 0};
  long long snap_cells = 0;
  // expanded region per detector for the adjoint scatter
  int elo[FDTDX_MAX_DETECTORS][3], ehi[FDTDX_MAX_DETECTORS][3];
  // workspace layout (byte offsets)
  size_t ws_snapH = 0, ws_snapE = 0, ws_partial = 0, ws_g = 0, ws_half = 0, ws_lpsi = 0, ws_bytes = 0;
  size_t ws_K = 0, ws_W = 0;  // full-tensor plans only
  size_t ws_Fo = 0, ws_ab[2] = {0, 0};  // full-tensor plans with conductivity
  size_t ws_fwd_end = 0;       // workspace bytes the two-sweep forward uses
  size_t g_off[FDTDX_MAX_DETECTORS], half_off[FDTDX_MAX_DETECTORS];
  // single-pass paths (1: LDS-tiled fused kernel, 2: register-march kernel): tile configuration
  // and ping-pong copies of E, H and psi in the workspace
  int cplx = 0;                // complex fields: every field array holds the real, then the imaginary part
  int use_fused = 0;
  FusedDyn fd;
  MarchDyn md;
  size_t fused_lds = 0;
  size_t ws_Ealt = 0, ws_Halt = 0, ws_psialt = 0;

  ~fdtdx_plan() {
    for (void* p : allocs) (void)hipFree(p);
  }
  template <class T>
  int upload(const T* host, size_t count, const T** out) {
    if (count == 0) { *out = nullptr; return FDTDX_OK; }
    if (!host) return fail(FDTDX_EINVAL, "missing host table");
    void* d = nullptr;
    HIP_TRY(hipMalloc(&d, count * sizeof(T)));
    allocs.push_back(d);
    HIP_TRY(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
    *out = (const T*)d;
    return FDTDX_OK;
  }
};

static bool box_ok(const int32_t* lo, const int32_t* hi, const StepStatic& P) {
  const int n[3] = {P.nx, P.ny, P.nz};
  for (int a = 0; a < 3; ++a)
    if (lo[a] < 0 || hi[a] > n[a] || lo[a] >= hi[a]) return false;
  return true;
}

extern "C" {

const char* fdtdx_last_error(void) { return g_err.c_str(); }

static int plan_create(const fdtdx_plan_desc* d, fdtdx_plan** out);

int fdtdx_plan_create(const fdtdx_plan_desc* d, fdtdx_plan** out) {
  std::lock_guard<std::mutex> lk(g_device_mu);
  return plan_create(d, out);
}

static int plan_create(const fdtdx_plan_desc* d, fdtdx_plan** out) {
  if (!d || !out) return fail(FDTDX_EINVAL, "null argument");
  *out = nullptr;
  if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return fail(FDTDX_EINVAL, "grid shape must be positive");
  if ((long long)d->ny * d->nz >= (1ll << 31)) return fail(FDTDX_ENOTSUP, "ny*nz too large");
  if (d->num_steps < 0) return fail(FDTDX_EINVAL, "num_steps < 0");
  if (d->eps_comps != 1 && d->eps_comps != 3 && d->eps_comps != 9)
    return fail(FDTDX_ENOTSUP, "eps_comps=%d: inverse permittivity is isotropic (1), diagonal (3) or a full tensor (9)",
                d->eps_comps);
  if (d->mu_comps != 0 && d->mu_comps != 1 && d->mu_comps != 3 && d->mu_comps != 9)
    return fail(FDTDX_ENOTSUP, "mu_comps=%d unsupported", d->mu_comps);
  if (d->num_cpml < 0 || d->num_cpml > FDTDX_MAX_CPML) return fail(FDTDX_EINVAL, "num_cpml out of range");
  if (d->num_sources < 0 || d->num_sources > FDTDX_MAX_SOURCES) return fail(FDTDX_EINVAL, "num_sources out of range");
  if (d->num_ade < 0 || d->num_ade > FDTDX_MAX_ADE) return fail(FDTDX_EINVAL, "num_ade out of range");
  if (d->num_detectors < 0 || d->num_detectors > FDTDX_MAX_DETECTORS)
    return fail(FDTDX_EINVAL, "num_detectors out of range");
  if (d->num_walls < 0 || d->num_walls > FDTDX_MAX_WALLS) return fail(FDTDX_EINVAL, "num_walls out of range");
  if ((d->num_cpml && !d->cpml) || (d->num_sources && !d->sources) || (d->num_ade && !d->ade) ||
      (d->num_detectors && !d->detectors) || (d->num_walls && !d->walls))
    return fail(FDTDX_EINVAL, "a descriptor array with a nonzero count is NULL");

  fdtdx_plan* pl = new fdtdx_plan();
  StepStatic& P = pl->P;
  memset(&P, 0, sizeof(P));
  P.nx = d->nx; P.ny = d->ny; P.nz = d->nz;
  P.T = d->num_steps;
  P.plane = (long long)d->ny * d->nz;
  P.N = P.plane * d->nx;
  P.c = d->courant;
  P.bpp = (int)((P.plane + kBlock - 1) / kBlock);
  P.nblocks = (long long)P.nx * P.bpp;
  P.nblocks_pad = pad_blocks(P.nblocks);
  // block order: XCD turns when a sweep's arrays (about 40 bytes per cell) exceed the 256 MiB
  // last-level cache, per-XCD ranges below that
  P.xorder = (double)P.N * 40.0 > 256.0 * 1024.0 * 1024.0 ? 1 : 0;
  if (const char* oenv = getenv("FDTDX_BLOCK_ORDER")) {  // "ranges" or "turns" (for measurements)
    if (strcmp(oenv, "ranges") == 0) P.xorder = 0;
    else if (strcmp(oenv, "turns") == 0) P.xorder = 1;
  }
  pl->ec = d->eps_comps;
  pl->mc = d->mu_comps;
  int rc;
  auto bail = [&](int code) { delete pl; return code; };

  // CPML slabs
  P.nslab = d->num_cpml;
  long long psi_off = 0;
  for (int s = 0; s < d->num_cpml; ++s) {
    const fdtdx_cpml_desc& c = d->cpml[s];
    SlabK& k = P.slab[s];
    const int n[3] = {P.nx, P.ny, P.nz};
    if (c.axis < 0 || c.axis > 2 || c.lo < 0 || c.hi > n[c.axis] || c.lo >= c.hi)
      return bail(fail(FDTDX_EINVAL, "cpml %d: bad range [%d,%d) on axis %d", s, c.lo, c.hi, c.axis));
    if (c.side != 0 && c.side != 1) return bail(fail(FDTDX_EINVAL, "cpml %d: side must be 0 or 1", s));
    k.axis = c.axis; k.lo = c.lo; k.hi = c.hi; k.L = c.hi - c.lo;
    k.sx = c.axis == 0 ? k.L : P.nx;
    k.sy = c.axis == 1 ? k.L : P.ny;
    k.sz = c.axis == 2 ? k.L : P.nz;
    k.cells = (long long)k.sx * k.sy * k.sz;
    k.kappa_one = c.kappa_is_one;
    k.psi_off = psi_off;
    psi_off += 4 * k.cells;
    k.iface = c.side == 0 ? c.hi - 1 : c.lo;
    const int o1 = c.axis == 0 ? 1 : 0, o2 = c.axis == 2 ? 1 : 2;
    k.pn0 = n[o1];
    k.pn1 = n[o2];
    k.rec_off_E = 3ll * k.pn0 * k.pn1;  // interface values of this slab's E layer (made bytes below)
    pl->iface_cells += (long long)k.pn0 * k.pn1;
    std::vector<float> cf(6 * k.L, 0.f);
    const float* src[6] = {c.b_E, c.a_E, c.ikm1_E, c.b_H, c.a_H, c.ikm1_H};
    for (int t = 0; t < 6; ++t) {
      if (!src[t]) {
        if ((t == 2 || t == 5) && c.kappa_is_one) continue;
        return bail(fail(FDTDX_EINVAL, "cpml %d: missing coefficient vector %d", s, t));
      }
      memcpy(&cf[t * k.L], src[t], sizeof(float) * k.L);
    }
    if ((rc = pl->upload(cf.data(), cf.size(), &k.coef))) return bail(rc);
  }
  {
    // interface record layout and time filter (recorder.py:32-68, time_filter.py:163-200)
    const fdtdx_record_desc& R = d->record;
    long long vals = 0, off = 0;
    for (int f = 0; f < 2; ++f)
      for (int s = 0; s < P.nslab; ++s) {
        SlabK& k = P.slab[s];
        const int fmt = f == 0 ? R.format_E[s] : R.format_H[s];
        if (fmt < FDTDX_REC_F32 || fmt > FDTDX_REC_E5M2FNUZ)
          return bail(fail(FDTDX_EINVAL, "record format %d of cpml %d is unknown", fmt, s));
        const long long n = 3ll * k.pn0 * k.pn1;
        const int w = fmt == FDTDX_REC_F32 ? 4 : (fmt <= FDTDX_REC_F16 ? 2 : 1);
        (f == 0 ? k.fmt_E : k.fmt_H) = fmt;
        (f == 0 ? k.rec_off_E : k.rec_off_H) = off;
        off = (off + n * w + 15) & ~15ll;
        vals += n;
      }
    P.rec_per_step = vals;
    P.rec_slot_bytes = off;
    P.rec_filter_first = R.filter_first ? 1 : 0;
    for (int s = 0; s < P.nslab; ++s) P.rec_narrow |= (P.slab[s].fmt_E | P.slab[s].fmt_H) != 0;
    const int T = P.T;
    const int kk = R.every_k > 1 ? R.every_k : 1;
    if (R.every_k < 0 || R.start < 0 || (kk > 1 && R.start >= T && T > 0))
      return bail(fail(FDTDX_EINVAL, "record time filter: every_k=%d start=%d for %d steps", R.every_k, R.start, T));
    pl->rec_slot_of.assign(T, -1);
    pl->rec_arr.assign(T, 0);
    pl->rec_factor.assign(T, 0.f);
    if (kk == 1 && R.start == 0) {
      for (int n = 0; n < T; ++n) pl->rec_slot_of[n] = pl->rec_arr[n] = n;
      pl->rec_slots = T;
    } else if (T > 0) {
      std::vector<int> save;
      for (int t = R.start; t < T; t += kk) save.push_back(t);
      if (save.empty() || save.back() != T - 1) save.push_back(T - 1);
      // step -> array index by the reference's forward fill: k - 1 rolls by one, the first k
      // entries reset to 0 after each
      std::vector<int> idx(T, 0), rolled(T);
      for (size_t i = 0; i < save.size(); ++i) idx[save[i]] = (int)i;
      for (int it = 0; it < kk - 1; ++it) {
        for (int t = 0; t < T; ++t) rolled[t] = idx[(t + T - 1) % T];
        for (int t = 0; t < T; ++t) if (idx[t] == 0) idx[t] = rolled[t];
        for (int t = 0; t < std::min(kk, T); ++t) idx[t] = 0;
      }
      auto first = [&](int v) { for (int t = 0; t < T; ++t) if (idx[t] == v) return t; return 0; };
      std::vector<char> kept(T, 0);
      for (int t : save) kept[t] = 1;
      for (int n = 0; n < T; ++n) {
        pl->rec_arr[n] = idx[n];
        if (kept[n]) { pl->rec_slot_of[n] = idx[n]; continue; }
        // (t - t_prev) / (t_next - t_prev), int32 differences divided in float32 (time_filter.py:239-243)
        const int tp = first(idx[n]), tn = first(idx[n] + 1);
        pl->rec_factor[n] = (float)(n - tp) / (float)(tn - tp);
      }
      pl->rec_slots = (long long)save.size();
      if (P.rec_per_step && (rc = pl->upload(pl->rec_slot_of.data(), (size_t)T, &P.rec_slot))) return bail(rc);
    }
    for (int n = 0; n < T; ++n)  // an interpolated step needs the slot after its own
      if (pl->rec_slot_of[n] < 0 && pl->rec_arr[n] + 1 >= pl->rec_slots)
        return bail(fail(FDTDX_EINVAL, "record time filter leaves step %d without a later slot", n));
  }
  pl->psi_floats = (size_t)psi_off;

  // sources
  P.nsrc = d->num_sources;
  for (int s = 0; s < d->num_sources; ++s) {
    const fdtdx_source_desc& c = d->sources[s];
    SrcK& k = P.src[s];
    k.kind = c.kind; k.field = c.field; k.comp = c.component;
    for (int a = 0; a < 3; ++a) { k.lo[a] = c.lo[a]; k.hi[a] = c.hi[a]; }
    if (c.field != 0 && c.field != 1) return bail(fail(FDTDX_EINVAL, "source %d: bad field", s));
    if (c.kind == FDTDX_SRC_POINT) {
      for (int a = 0; a < 3; ++a) k.hi[a] = c.lo[a] + 1;
      if (!box_ok(k.lo, k.hi, P)) return bail(fail(FDTDX_EINVAL, "source %d: cell outside the grid", s));
      if (c.component < 0 || c.component > 2) return bail(fail(FDTDX_EINVAL, "source %d: bad component", s));
      if ((rc = pl->upload(c.table, (size_t)P.T, &k.table))) return bail(rc);
    } else if (c.kind == FDTDX_SRC_PLANE) {
      if (!box_ok(k.lo, k.hi, P)) return bail(fail(FDTDX_EINVAL, "source %d: plane box outside the grid", s));
      k.py = k.hi[1] - k.lo[1];
      k.pz = k.hi[2] - k.lo[2];
      k.pcells = (long long)(k.hi[0] - k.lo[0]) * k.py * k.pz;
      for (int t = 0; t < 2; ++t) {
        k.target[t] = c.target[t]; k.src[t] = c.src_comp[t]; k.sign[t] = c.sign[t];
        if (c.target[t] < 0 || c.target[t] > 2 || c.src_comp[t] < 0 || c.src_comp[t] > 2)
          return bail(fail(FDTDX_EINVAL, "source %d: bad plane term", s));
      }
      if ((rc = pl->upload(c.table, 3 * (size_t)P.T, &k.table))) return bail(rc);
      if ((rc = pl->upload(c.profile, 3 * (size_t)k.pcells, &k.profile))) return bail(rc);
    } else {
      return bail(fail(FDTDX_ENOTSUP, "source %d: unknown kind %d", s, c.kind));
    }
  }

  // PEC / PMC faces
  P.nwall = d->num_walls;
  for (int w = 0; w < d->num_walls; ++w) {
    const fdtdx_wall_desc& c = d->walls[w];
    const int n[3] = {P.nx, P.ny, P.nz};
    if (c.kind != FDTDX_WALL_PEC && c.kind != FDTDX_WALL_PMC) return bail(fail(FDTDX_EINVAL, "wall %d: kind", w));
    if (c.axis < 0 || c.axis > 2 || c.index < 0 || c.index >= n[c.axis])
      return bail(fail(FDTDX_EINVAL, "wall %d: layer outside the grid", w));
    P.wall[w].kind = c.kind; P.wall[w].axis = c.axis; P.wall[w].index = c.index;
  }

  // periodic axes: the curls' neighbour on the far face wraps to the near face
  for (int a = 0; a < 3; ++a) {
    if (d->periodic[a] != 0 && d->periodic[a] != 1) return bail(fail(FDTDX_EINVAL, "periodic[%d] must be 0 or 1", a));
    P.wrap[a] = d->periodic[a];
    if (d->mirror[a] != 0 && d->mirror[a] != 1) return bail(fail(FDTDX_EINVAL, "mirror[%d] must be 0 or 1", a));
    if (d->mirror[a] && d->periodic[a])
      return bail(fail(FDTDX_EINVAL, "axis %d: a symmetry plane and a periodic axis exclude each other", a));
    P.mirror[a] = d->mirror[a];
  }
  // complex fields: the Bloch phase of each periodic face
  if (d->complex_fields != 0 && d->complex_fields != 1) return bail(fail(FDTDX_EINVAL, "complex_fields must be 0 or 1"));
  pl->cplx = d->complex_fields;
  P.bloch = 0;
  for (int a = 0; a < 3; ++a)
    for (int side = 0; side < 2; ++side) {
      const double th = pl->cplx ? d->bloch_phase[a][side] : 0.0;
      if (!std::isfinite(th)) return bail(fail(FDTDX_EINVAL, "bloch_phase[%d][%d] is not finite", a, side));
      if (th != 0.0 && !P.wrap[a])
        return bail(fail(FDTDX_EINVAL, "bloch_phase[%d][%d] is nonzero on an axis that is not periodic", a, side));
      P.bcos[a][side] = (float)std::cos(th);
      P.bsin[a][side] = (float)std::sin(th);
      if (th != 0.0) P.bloch = 1;
    }
  for (int s = 0; s < P.nslab; ++s)
    if (P.wrap[P.slab[s].axis]) return bail(fail(FDTDX_EINVAL, "cpml %d lies on periodic axis %d", s, P.slab[s].axis));
  for (int w = 0; w < P.nwall; ++w)
    if (P.wrap[P.wall[w].axis]) return bail(fail(FDTDX_EINVAL, "wall %d lies on periodic axis %d", w, P.wall[w].axis));

  // conductivity boxes
  P.eta0 = (float)(4e-7 * 3.141592653589793 * 299792458.0);  // constants.py:16-22
  for (int w = 0; w < 2; ++w) {
    const fdtdx_loss_desc& c = w == 0 ? d->sigma_E : d->sigma_H;
    LossK& k = P.loss[w];
    const char* nm = w == 0 ? "sigma_E" : "sigma_H";
    if (c.comps == 0) continue;
    if (c.comps != 1 && c.comps != 3 && c.comps != 9) return bail(fail(FDTDX_EINVAL, "%s: comps must be 0, 1, 3 or 9", nm));
    if (c.comps == 9 && (w == 0 ? d->eps_comps : d->mu_comps) != 9)
      return bail(fail(FDTDX_EINVAL, "%s: a full conductivity tensor needs the full-tensor %s (9 components)", nm,
                       w == 0 ? "inverse permittivity" : "inverse permeability"));
    for (int a = 0; a < 3; ++a) { k.lo[a] = c.lo[a]; k.hi[a] = c.hi[a]; }
    if (!box_ok(k.lo, k.hi, P)) return bail(fail(FDTDX_EINVAL, "%s: box outside the grid", nm));
    if (!c.sigma) return bail(fail(FDTDX_EINVAL, "%s: sigma is NULL", nm));
    k.comps = c.comps;
    k.by = k.hi[1] - k.lo[1]; k.bz = k.hi[2] - k.lo[2];
    k.cells = (long long)(k.hi[0] - k.lo[0]) * k.by * k.bz;
    if ((rc = pl->upload(c.sigma, (size_t)k.comps * k.cells, &k.sig))) return bail(rc);
  }

  // ADE boxes
  P.nade = d->num_ade;
  long long ade_off = 0;
  for (int b = 0; b < d->num_ade; ++b) {
    const fdtdx_ade_desc& c = d->ade[b];
    AdeK& k = P.ade[b];
    for (int a = 0; a < 3; ++a) { k.lo[a] = c.lo[a]; k.hi[a] = c.hi[a]; }
    if (!box_ok(k.lo, k.hi, P)) return bail(fail(FDTDX_EINVAL, "ade %d: box outside the grid", b));
    if (c.num_poles < 1 || c.num_poles > FDTDX_MAX_POLES) return bail(fail(FDTDX_ENOTSUP, "ade %d: poles", b));
    if (c.coef_comps != 1 && c.coef_comps != 3) return bail(fail(FDTDX_EINVAL, "ade %d: coef_comps", b));
    k.poles = c.num_poles; k.ccomp = c.coef_comps;
    k.c3comp = c.coupling_comps ? c.coupling_comps : c.coef_comps;
    if (k.c3comp != k.ccomp && !(k.c3comp == 9 && d->eps_comps == 9))
      return bail(fail(FDTDX_EINVAL, "ade %d: coupling_comps must be 0, coef_comps, or 9 with eps_comps 9", b));
    k.by = k.hi[1] - k.lo[1]; k.bz = k.hi[2] - k.lo[2];
    k.cells = (long long)(k.hi[0] - k.lo[0]) * k.by * k.bz;
    k.state_off = ade_off;
    ade_off += 3ll * k.poles * 3 * k.cells;
    const size_t per = (size_t)k.poles * k.ccomp * k.cells;
    const size_t per3 = (size_t)k.poles * k.c3comp * k.cells;
    std::vector<float> cf(2 * per + per3);
    if (!c.c1 || !c.c2 || !c.c3) return bail(fail(FDTDX_EINVAL, "ade %d: missing coefficients", b));
    memcpy(&cf[0], c.c1, per * 4);
    memcpy(&cf[per], c.c2, per * 4);
    memcpy(&cf[2 * per], c.c3, per3 * 4);
    if ((rc = pl->upload(cf.data(), cf.size(), &k.coef))) return bail(rc);
  }
  pl->ade_floats = (size_t)ade_off;

  // detectors
  pl->ndet = d->num_detectors;
  pl->tti.resize(pl->ndet);
  pl->window.resize(pl->ndet);
  int ulo[3] = {P.nx, P.ny, P.nz}, uhi[3] = {0, 0, 0};
  for (int q = 0; q < pl->ndet; ++q) {
    const fdtdx_detector_desc& c = d->detectors[q];
    DetK& k = pl->det[q];
    memset(&k, 0, sizeof(k));
    k.kind = c.kind;
    for (int a = 0; a < 3; ++a) { k.lo[a] = c.lo[a]; k.hi[a] = c.hi[a]; }
    if (!box_ok(k.lo, k.hi, P)) return bail(fail(FDTDX_EINVAL, "detector %d: box outside the grid", q));
    if (c.kind < 0 || c.kind > 3) return bail(fail(FDTDX_ENOTSUP, "detector %d: kind %d", q, c.kind));
    k.reduce = c.reduce_volume; k.exact = c.exact_interpolation;
    if (k.reduce < 0 || k.reduce > (c.kind == FDTDX_DET_ENERGY ? 2 : 1) ||
        (k.reduce == 2 && (c.axis < 0 || c.axis > 2)))
      return bail(fail(FDTDX_EINVAL, "detector %d: reduce_volume %d", q, k.reduce));
    k.ncomp = c.num_components;
    if (k.ncomp < 0 || k.ncomp > 6) return bail(fail(FDTDX_EINVAL, "detector %d: components", q));
    for (int t = 0; t < k.ncomp; ++t) {
      k.comps[t] = c.components[t];
      if (k.comps[t] < 0 || k.comps[t] > 5) return bail(fail(FDTDX_EINVAL, "detector %d: component id", q));
    }
    k.cplx = pl->cplx;
    if (k.cplx && k.kind == FDTDX_DET_PHASOR && k.reduce && 2 * k.ncomp > kDetMaxValues)
      return bail(fail(FDTDX_ENOTSUP, "detector %d: a reduced phasor of complex fields records at most %d components",
                       q, kDetMaxValues / 2));
    k.axis = c.axis; k.sign = c.sign; k.weight = c.weight; k.nfreq = c.num_freqs; k.scale = c.scale;
    k.cells = (long long)(k.hi[0] - k.lo[0]) * (k.hi[1] - k.lo[1]) * (k.hi[2] - k.lo[2]);
    if (!c.time_to_idx) return bail(fail(FDTDX_EINVAL, "detector %d: missing time_to_idx", q));
    pl->tti[q].assign(c.time_to_idx, c.time_to_idx + P.T);
    k.accum = c.accumulate ? 1 : 0;
    k.inv = c.inverse ? 1 : 0;
    if (k.inv && k.accum) return bail(fail(FDTDX_EINVAL, "detector %d: an inverse detector cannot accumulate", q));
    if (k.accum && !(q > 0 && k.reduce && (k.kind == FDTDX_DET_ENERGY || k.kind == FDTDX_DET_FLUX) &&
                     pl->det[q - 1].kind == k.kind && pl->det[q - 1].reduce &&
                     reduced_values(pl->det[q - 1]) == reduced_values(k) && pl->tti[q - 1] == pl->tti[q]))
      return bail(fail(FDTDX_EINVAL, "detector %d: accumulate needs a reducing energy or flux detector after one "
                                     "of the same kind, value count and time_to_idx", q));
    if ((rc = pl->upload((const int*)c.time_to_idx, (size_t)P.T, &k.tti))) return bail(rc);
    if (c.kind == FDTDX_DET_PHASOR) {
      if (k.nfreq < 1) return bail(fail(FDTDX_EINVAL, "detector %d: no frequencies", q));
      if ((rc = pl->upload(c.cos_table, (size_t)k.nfreq * P.T, &k.cos_t))) return bail(rc);
      if ((rc = pl->upload(c.sin_table, (size_t)k.nfreq * P.T, &k.sin_t))) return bail(rc);
      if ((rc = pl->upload(c.window, (size_t)P.T, &k.window))) return bail(rc);
      pl->window[q].assign(c.window, c.window + P.T);
    }
    const int n[3] = {P.nx, P.ny, P.nz};
    for (int a = 0; a < 3; ++a)  // the snapshot box holds every wrapped neighbour only in these two cases
      if (P.wrap[a] && !(c.lo[a] == 0 && c.hi[a] == n[a]) && (c.lo[a] < 1 || c.hi[a] > n[a] - 1))
        return bail(fail(FDTDX_EINVAL, "detector %d: on periodic axis %d it must span the axis or stay one cell "
                                       "inside both faces", q, a));
    for (int a = 0; a < 3; ++a) {
      pl->elo[q][a] = std::max(0, c.lo[a] - 1);
      pl->ehi[q][a] = std::min(n[a], c.hi[a] + 1);
      ulo[a] = std::min(ulo[a], pl->elo[q][a]);
      uhi[a] = std::max(uhi[a], pl->ehi[q][a]);
    }
  }
  if (pl->ndet > 0) {
    pl->has_snap = 1;
    for (int a = 0; a < 3; ++a) { pl->slo[a] = ulo[a]; pl->shi[a] = uhi[a]; }
    pl->snap_cells = (long long)(uhi[0] - ulo[0]) * (uhi[1] - ulo[1]) * (uhi[2] - ulo[2]);
  }
  if (pl->ndet > 0 && (rc = pl->upload(pl->det, (size_t)pl->ndet, &pl->det_dev))) return bail(rc);
  // Tile masks: for each logical tile (kBlock consecutive cells of one x-plane) the descriptors
  // whose box meets the tile's footprint, so the sweeps skip the others without per-cell tests.
  // Two sets: row-major tiles (reverse and adjoint sweeps) and tiles of the forward sweeps'
  // permuted plane order (z-interior cells first, then the z-slab cells).
  {
    P.zlo = 0;
    P.zhi = P.nz;
    for (int q = 0; q < P.nslab; ++q) {
      const SlabK& sl = P.slab[q];
      if (sl.axis != 2) continue;
      if (sl.lo == 0) P.zlo = std::max(P.zlo, sl.hi);
      if (sl.hi == P.nz) P.zhi = std::min(P.zhi, sl.lo);
    }
    if (P.zlo >= P.zhi) { P.zlo = 0; P.zhi = P.nz; }
    // mask of the descriptors meeting the inclusive footprint {i} x [jlo, jhi] x [klo, khi];
    // `wide`: slabs, walls and conductivity boxes count one cell wider (the reverse and adjoint
    // sweeps read their neighbours' terms), else exactly (the forward sweeps)
    auto box_mask = [&](int i, int jlo, int jhi, int klo, int khi, bool wide) -> uint32_t {
      const int g = wide ? 1 : 0;
      const int flo[3] = {i, jlo, klo}, fhi[3] = {i, jhi, khi};
      auto meets = [&](const int* lo, const int* hi) {
        for (int a = 0; a < 3; ++a)
          if (fhi[a] < lo[a] || flo[a] >= hi[a]) return false;
        return true;
      };
      uint32_t m = 0;
      for (int q = 0; q < P.nslab; ++q) {
        const SlabK& sl = P.slab[q];
        if (fhi[sl.axis] + g >= sl.lo && flo[sl.axis] - g < sl.hi) m |= 1u << q;
      }
      for (int q = 0; q < P.nsrc; ++q)
        if (meets(P.src[q].lo, P.src[q].hi)) m |= 1u << (kMaskSrc + q);
      for (int q = 0; q < P.nade; ++q)
        if (meets(P.ade[q].lo, P.ade[q].hi)) m |= 1u << (kMaskAde + q);
      if (pl->has_snap && meets(pl->slo, pl->shi)) m |= 1u << kMaskSnap;
      for (int q = 0; q < P.nwall; ++q) {
        const WallK& w = P.wall[q];
        if (fhi[w.axis] + g >= w.index && flo[w.axis] - g <= w.index) m |= 1u << (kMaskWall + q);
      }
      // conductivity boxes: when wide, one cell wider on every axis, wrapping across periodic faces
      for (int q = 0; q < 2; ++q) {
        const LossK& L = P.loss[q];
        if (!L.comps) continue;
        const int n[3] = {P.nx, P.ny, P.nz};
        bool hit = true;
        for (int a = 0; a < 3 && hit; ++a) {
          const int lo = flo[a] - g, hi = fhi[a] + g;  // inclusive
          bool ax = hi >= L.lo[a] && lo < L.hi[a];
          if (P.wrap[a]) ax = ax || (lo < 0 && L.hi[a] > lo + n[a]) || (hi >= n[a] && L.lo[a] <= hi - n[a]);
          hit = ax;
        }
        if (hit) m |= 1u << (kMaskLoss + q);
      }
      return m;
    };
    // row-major span [p0, p1] of a plane of rows of `w` cells: its footprint
    auto span_mask = [&](int i, long long p0, long long p1, int w, bool wide) -> uint32_t {
      const int jlo = (int)(p0 / w), jhi = (int)(p1 / w);
      const int klo = jlo == jhi ? (int)(p0 % w) : 0;
      const int khi = jlo == jhi ? (int)(p1 % w) : w - 1;
      return box_mask(i, jlo, jhi, klo, khi, wide);
    };
    const bool perm = P.zlo > 0 || P.zhi < P.nz;
    // rows per group: the smallest count whose z-interior and z-slab parts are whole tiles, if
    // one up to 64 exists, else 32
    P.zg = P.ny;
    if (perm) {
      const int W = P.zhi - P.zlo, Wz = P.nz - W;
      P.zg = 32;
      for (int g = 1; g <= 64; ++g)
        if ((g * W) % kBlock == 0 && (g * Wz) % kBlock == 0) { P.zg = g; break; }
      P.zg = std::min(P.zg, P.ny);
    }
    std::vector<uint32_t> tm((size_t)P.nblocks, 0u), fm((size_t)P.nblocks, 0u);
    for (long long lb = 0; lb < P.nblocks; ++lb) {
      const int i = (int)(lb / P.bpp);
      const long long p0 = (lb - (long long)i * P.bpp) * kBlock;
      const long long p1 = std::min(p0 + kBlock, P.plane) - 1;
      tm[(size_t)lb] = span_mask(i, p0, p1, P.nz, true);
      if (!perm) {
        fm[(size_t)lb] = span_mask(i, p0, p1, P.nz, false);
        continue;
      }
      // footprints of the tile's z-interior and z-slab cells, from the sweeps' own mapping
      int lo[2][2] = {{P.ny, P.nz}, {P.ny, P.nz}}, hi[2][2] = {{-1, -1}, {-1, -1}};
      for (long long p = p0; p <= p1; ++p) {
        int j, k;
        perm_cell(P, (int)p, j, k);
        const int c = (k < P.zlo || k >= P.zhi) ? 1 : 0;
        lo[c][0] = std::min(lo[c][0], j); hi[c][0] = std::max(hi[c][0], j);
        lo[c][1] = std::min(lo[c][1], k); hi[c][1] = std::max(hi[c][1], k);
      }
      uint32_t m = 0;
      for (int c = 0; c < 2; ++c)
        if (hi[c][0] >= 0) m |= box_mask(i, lo[c][0], hi[c][0], lo[c][1], hi[c][1], false);
      fm[(size_t)lb] = m;
    }
    if ((rc = pl->upload(tm.data(), tm.size(), &P.tile_mask))) return bail(rc);
    if ((rc = pl->upload(fm.data(), fm.size(), &P.fwd_mask))) return bail(rc);
    // quad sweeps: 4 z cells per lane
    P.vq = 0;
    P.cax = 0;
    P.acax = 0;
    const char* qenv = getenv("FDTDX_SWEEP");
    const bool quad_ok = P.nz % 4 == 0 && d->eps_comps <= 3 && d->mu_comps <= 3 && !pl->cplx &&
                         !(qenv && strcmp(qenv, "quad") != 0 && strcmp(qenv, "default") != 0);
    // per-axis CPML: every slab spans its face, at most one slab per side of an axis, the two sides
    // apart, and the x slabs before the y slabs before the z slabs in object order (the order in
    // which the corrections are added). The one-cell adjoint sweeps use it as it is (acax); the
    // quad sweeps (cax) also need no quad meeting both z slabs.
    {
      bool ok = true;
      int lo_id[3] = {99, 99, 99}, hi_id[3] = {-1, -1, -1};
      for (int a = 0; a < 3; ++a) P.axs[a][0] = P.axs[a][1] = -1;
      for (int q = 0; q < P.nslab && ok; ++q) {
        const SlabK& sl = P.slab[q];
        const int a = sl.axis, n_a = a == 0 ? P.nx : (a == 1 ? P.ny : P.nz);
        const int side = sl.lo == 0 ? 0 : (sl.hi == n_a ? 1 : -1);
        if (side < 0 || P.axs[a][side] >= 0) { ok = false; break; }
        P.axs[a][side] = q;
        lo_id[a] = std::min(lo_id[a], q);
        hi_id[a] = std::max(hi_id[a], q);
      }
      for (int a = 0; a < 3 && ok; ++a)
        if (P.axs[a][0] >= 0 && P.axs[a][1] >= 0 && P.slab[P.axs[a][0]].hi > P.slab[P.axs[a][1]].lo) ok = false;
      int ord[3] = {0, 1, 2};
      std::stable_sort(ord, ord + 3, [&](int x, int y) { return lo_id[x] < lo_id[y]; });
      bool xyz = ok;  // of two axes with slabs, every slab of the lower axis comes first
      for (int a = 0; a < 3; ++a)
        for (int b2 = a + 1; b2 < 3; ++b2)
          if (hi_id[a] >= 0 && hi_id[b2] >= 0 && hi_id[a] > lo_id[b2]) xyz = false;
      // FDTDX_REVERSE_CPML=loop keeps the reverse and adjoint sweeps on the slab loop (the same
      // results; for comparing the two paths)
      const char* renv = getenv("FDTDX_REVERSE_CPML");
      P.acax = xyz && !(renv && strcmp(renv, "loop") == 0) ? 1 : 0;
      const bool zq = !(P.axs[2][0] >= 0 && P.axs[2][1] >= 0 &&
                        (P.slab[P.axs[2][0]].hi - 1) / 4 >= P.slab[P.axs[2][1]].lo / 4);  // no quad meets both
      P.cax = quad_ok && xyz && zq ? 1 : 0;
      for (int u = 0; u < 3; ++u) P.axord[u] = ord[u];
      for (int a = 0; a < 3; ++a)
        for (int side = 0; side < 2; ++side) {
          StepStatic::AxSlab& x = P.axd[a][side];
          const int q = P.axs[a][side];
          if (q < 0) {  // no slab: an empty range
            x.lo = side ? (a == 0 ? P.nx : (a == 1 ? P.ny : P.nz)) : 0;
            x.hi = x.lo; x.L = 0; x.kappa_one = 1; x.cells = 0; x.psi_off = 0; x.coef = nullptr; x.aligned = 0;
            continue;
          }
          const SlabK& sl = P.slab[q];
          x.lo = sl.lo; x.hi = sl.hi; x.L = sl.L; x.kappa_one = sl.kappa_one;
          x.aligned = (sl.lo % 4 == 0 && sl.L % 4 == 0) ? 1 : 0;
          x.cells = sl.cells; x.psi_off = sl.psi_off; x.coef = sl.coef;
        }
    }
    if (quad_ok) {
      // Without the per-axis CPML the quads of a plane run in the row-grouped order (z-interior
      // quads of a group of rows, then its z-slab quads) when the z-slab edges fall on quad
      // boundaries. With it they run row-major: the 400x400x200 forward measured 0.293 / 0.261 ms
      // per E / H sweep row-major against 0.313 / 0.283 ms in the row-grouped order.
      const int nzq = P.nz / 4;
      P.vzlo = 0;
      P.vzhi = nzq;
      if (!P.cax && perm && P.zlo % 4 == 0 && P.zhi % 4 == 0) {
        P.vzlo = P.zlo / 4;
        P.vzhi = P.zhi / 4;
      }
      const bool vperm = P.vzlo > 0 || P.vzhi < nzq;
      P.vzg = P.ny;
      if (vperm) {
        const int W = P.vzhi - P.vzlo, Wz = nzq - W;
        P.vzg = 32;
        for (int g = 1; g <= 64; ++g)
          if ((g * W) % kBlock == 0 && (g * Wz) % kBlock == 0) { P.vzg = g; break; }
        P.vzg = std::min(P.vzg, P.ny);
      }
      const long long nq = (long long)P.ny * nzq;
      P.vbpp = (int)((nq + kBlock - 1) / kBlock);
      P.vnblocks = (long long)P.nx * P.vbpp;
      P.vnblocks_pad = pad_blocks(P.vnblocks);
      std::vector<uint32_t> vm((size_t)P.vnblocks, 0u);
      for (long long lb = 0; lb < P.vnblocks; ++lb) {
        const int i = (int)(lb / P.vbpp);
        const long long q0 = (lb - (long long)i * P.vbpp) * kBlock;
        const long long q1 = std::min(q0 + kBlock, nq) - 1;
        int lo[2][2] = {{P.ny, P.nz}, {P.ny, P.nz}}, hi[2][2] = {{-1, -1}, {-1, -1}};
        for (long long q = q0; q <= q1; ++q) {
          int j, k;
          perm_quad(P, (int)q, j, k);
          const int c = (k < P.zlo || k >= P.zhi) ? 1 : 0;
          lo[c][0] = std::min(lo[c][0], j); hi[c][0] = std::max(hi[c][0], j);
          lo[c][1] = std::min(lo[c][1], k); hi[c][1] = std::max(hi[c][1], k + 3);
        }
        uint32_t m = 0;
        for (int c = 0; c < 2; ++c)
          if (hi[c][0] >= 0) m |= box_mask(i, lo[c][0], hi[c][0], lo[c][1], hi[c][1], false);
        vm[(size_t)lb] = m;
      }
      if ((rc = pl->upload(vm.data(), vm.size(), &P.vec_mask))) return bail(rc);
      // listed tiles, one cell per thread: those of the E half-step have descriptors beyond CPML
      // slabs and the snapshot; those of the H half-step beyond these and ADE boxes, which only
      // the E half-step reads (a quad-vectorised ADE sweep measured slower than the one-cell one:
      // 0.245 against 0.236 ms per E sweep of the 256^3 dispersive case)
      // The descriptors each half-step reads: E sources, PEC faces, ADE boxes and electric
      // conductivity for E; H sources, PMC faces and magnetic conductivity for H.
      uint32_t eb = ((1u << FDTDX_MAX_ADE) - 1u) << kMaskAde | (1u << kMaskLoss);
      uint32_t hb = 1u << (kMaskLoss + 1);
      for (int q = 0; q < P.nsrc; ++q) (P.src[q].field == 0 ? eb : hb) |= 1u << (kMaskSrc + q);
      for (int w = 0; w < P.nwall; ++w) (P.wall[w].kind == FDTDX_WALL_PEC ? eb : hb) |= 1u << (kMaskWall + w);
      P.vlistE = eb;
      P.vlistH = hb;
      std::vector<int> slow, slowH;
      for (long long lb = 0; lb < P.vnblocks; ++lb) {
        if (vm[(size_t)lb] & eb) slow.push_back((int)lb);
        if (vm[(size_t)lb] & hb) slowH.push_back((int)lb);
      }
      P.vnslow = (long long)slow.size();
      P.vslow = nullptr;
      if (!slow.empty() && (rc = pl->upload(slow.data(), slow.size(), &P.vslow))) return bail(rc);
      P.vnslowH = (long long)slowH.size();
      P.vslowH = nullptr;
      if (!slowH.empty() && (rc = pl->upload(slowH.data(), slowH.size(), &P.vslowH))) return bail(rc);
      P.vq = 1;
    }
  }

  // full tensors (eps_comps or mu_comps 9): the tensor sweeps carry no conductivity, poles, plane
  // sources or energy detectors
  const bool tensor = pl->ec == 9 || pl->mc == 9;
  if (tensor && pl->cplx) return bail(fail(FDTDX_ENOTSUP, "complex fields in full-tensor media are not supported"));
  if (pl->cplx)
    for (int q = 0; q < pl->ndet; ++q)
      if (pl->det[q].inv) return bail(fail(FDTDX_ENOTSUP, "detector %d: inverse detectors need the reverse run, "
                                                          "which complex-field plans do not have", q));
  // conductivity of a full-tensor half-step: per-cell update matrices inside its box
  P.tsig[0] = pl->ec == 9 && P.loss[0].comps ? 1 : 0;
  P.tsig[1] = pl->mc == 9 && P.loss[1].comps ? 1 : 0;
  if (tensor) {
    for (int s = 0; s < P.nsrc; ++s)
      if (P.src[s].kind == FDTDX_SRC_PLANE)
        return bail(fail(FDTDX_ENOTSUP, "source %d: plane sources in full-tensor media are not supported", s));
    for (int q = 0; q < pl->ndet; ++q)
      if (pl->det[q].kind == FDTDX_DET_ENERGY)
        return bail(fail(FDTDX_ENOTSUP, "detector %d: energy detectors in full-tensor media are not supported", q));
  }

  // workspace
  size_t off = 0;
  if (tensor) {  // the curl of the current half-step, and the adjoint's transposed-tensor cotangent
    pl->ws_K = off; off = align256(off + 3 * (size_t)P.N * 4);
    pl->ws_W = off; off = align256(off + 3 * (size_t)P.N * 4);
  }
  if (P.tsig[0] || P.tsig[1]) {  // the pre-update field and the conductivity boxes' matrices
    pl->ws_Fo = off; off = align256(off + 3 * (size_t)P.N * 4);
    for (int w = 0; w < 2; ++w)
      if (P.tsig[w]) { pl->ws_ab[w] = off; off = align256(off + 36 * (size_t)P.loss[w].cells * 4); }
  }
  pl->ws_snapH = off; off = align256(off + 3 * (size_t)pl->snap_cells * 4);
  pl->ws_snapE = off; off = align256(off + 3 * (size_t)pl->snap_cells * 4);
  pl->ws_partial = off; off = align256(off + (size_t)pl->ndet * kDetMaxValues * kDetMaxBlocks * 8);
  pl->ws_fwd_end = off;  // the forward (two-sweep schedules) touches nothing past this offset
  for (int q = 0; q < pl->ndet; ++q) { pl->g_off[q] = off; off = align256(off + 6 * (size_t)pl->det[q].cells * 4); }
  for (int q = 0; q < pl->ndet; ++q) {
    const long long ec = (long long)(pl->ehi[q][0] - pl->elo[q][0]) * (pl->ehi[q][1] - pl->elo[q][1]) *
                         (pl->ehi[q][2] - pl->elo[q][2]);
    pl->half_off[q] = off;
    off = align256(off + 3 * (size_t)ec * 4);
  }
  pl->ws_lpsi = off; off = align256(off + pl->psi_floats * 4);
  {
    // Sweep schedule, read at plan creation from FDTDX_SWEEP: "two" (default) runs the E and H
    // sweeps in place; "fused" (LDS tiles) and "march" (register march, FDTDX_MARCH_TY rows and
    // FDTDX_MARCH_XC planes per wavefront) are single-pass and ping-pong E, H and psi.
    const char* mode = getenv("FDTDX_SWEEP");
    memset(&pl->fd, 0, sizeof(pl->fd));
    memset(&pl->md, 0, sizeof(pl->md));
    bool alt = false;
    // PEC / PMC faces, periodic wrap, conductivity and narrow record formats are handled by the
    // two-sweep kernels only
    bool narrow = false;
    for (int s = 0; s < P.nslab; ++s) narrow = narrow || P.slab[s].fmt_E || P.slab[s].fmt_H;
    if (P.nwall > 0 || P.wrap[0] || P.wrap[1] || P.wrap[2] || P.loss[0].comps || P.loss[1].comps || narrow || tensor ||
        pl->cplx)
      mode = nullptr;
    if (mode && strcmp(mode, "fused") == 0) {
      const char* tgt = getenv("FDTDX_FUSED_TARGET");
      const int target = tgt ? atoi(tgt) : 512;
      if (fused_config(P, target > 0 ? target : 512, pl->fd, pl->fused_lds)) { pl->use_fused = 1; alt = true; }
    } else if (mode && strcmp(mode, "march") == 0) {
      const char* ty = getenv("FDTDX_MARCH_TY");
      const char* xc = getenv("FDTDX_MARCH_XC");
      if (march_config(P, ty ? atoi(ty) : 2, xc ? atoi(xc) : 25, pl->md)) { pl->use_fused = 2; alt = true; }
    }
    if (alt) {
      pl->ws_Ealt = off; off = align256(off + 3 * (size_t)P.N * 4);
      pl->ws_Halt = off; off = align256(off + 3 * (size_t)P.N * 4);
      pl->ws_psialt = off; off = align256(off + pl->psi_floats * 4);
    }
  }
  pl->ws_bytes = off;
  *out = pl;
  return FDTDX_OK;
}

void fdtdx_plan_destroy(fdtdx_plan* plan) {
  std::lock_guard<std::mutex> lk(g_device_mu);
  delete plan;
}

int fdtdx_psi_floats(const fdtdx_plan* p, size_t* out) {
  if (!p || !out) return fail(FDTDX_EINVAL, "null argument");
  *out = p->psi_floats * (p->cplx ? 2 : 1);
  return FDTDX_OK;
}
int fdtdx_ade_floats(const fdtdx_plan* p, size_t* out) {
  if (!p || !out) return fail(FDTDX_EINVAL, "null argument");
  *out = p->ade_floats * (p->cplx ? 2 : 1);
  return FDTDX_OK;
}
int fdtdx_record_slot_bytes(const fdtdx_plan* p, size_t* out) {
  if (!p || !out) return fail(FDTDX_EINVAL, "null argument");
  *out = (size_t)p->P.rec_slot_bytes;
  return FDTDX_OK;
}

int fdtdx_record_slots(const fdtdx_plan* p, size_t* out) {
  if (!p || !out) return fail(FDTDX_EINVAL, "null argument");
  *out = (size_t)p->rec_slots;
  return FDTDX_OK;
}

int fdtdx_record_slot_of(const fdtdx_plan* p, int32_t step, int32_t* out) {
  if (!p || !out) return fail(FDTDX_EINVAL, "null argument");
  if (step < 0 || step >= p->P.T) return fail(FDTDX_EINVAL, "step %d outside [0,%d)", step, p->P.T);
  *out = p->rec_slot_of[step];
  return FDTDX_OK;
}

int fdtdx_record_floats_per_step(const fdtdx_plan* p, size_t* out) {
  if (!p || !out) return fail(FDTDX_EINVAL, "null argument");
  *out = (size_t)p->P.rec_per_step;
  return FDTDX_OK;
}
int fdtdx_workspace_bytes(const fdtdx_plan* p, size_t* out) {
  if (!p || !out) return fail(FDTDX_EINVAL, "null argument");
  *out = p->ws_bytes;
  return FDTDX_OK;
}

static int check_buffers(const fdtdx_plan* p, const fdtdx_buffers* b) {
  if (!b || !b->E || !b->H || !b->inv_eps) return fail(FDTDX_EINVAL, "E, H and inv_eps are required");
  if (p->mc && !b->inv_mu) return fail(FDTDX_EINVAL, "plan has mu_comps=%d but inv_mu is NULL", p->mc);
  if (p->psi_floats && !b->psi) return fail(FDTDX_EINVAL, "plan has CPML slabs but psi is NULL");
  if (p->ade_floats && !b->ade_state) return fail(FDTDX_EINVAL, "plan has ADE boxes but ade_state is NULL");
  if (p->ws_bytes && !b->workspace) return fail(FDTDX_EINVAL, "workspace is NULL");
  for (int q = 0; q < p->ndet; ++q)
    if (!b->detector_state || !b->detector_state[q]) return fail(FDTDX_EINVAL, "detector %d state is NULL", q);
  return FDTDX_OK;
}

static StepDyn make_dyn(const fdtdx_plan* p, const fdtdx_buffers* b) {
  StepDyn D;
  memset(&D, 0, sizeof(D));
  D.E = b->E; D.H = b->H; D.ie = b->inv_eps; D.im = b->inv_mu;
  D.psi = b->psi; D.ade = b->ade_state;
  if (p->ec == 9 || p->mc == 9) D.K = (float*)((char*)b->workspace + p->ws_K);
  if (p->P.tsig[0] || p->P.tsig[1]) {
    D.Fo = (float*)((char*)b->workspace + p->ws_Fo);
    for (int w = 0; w < 2; ++w)
      if (p->P.tsig[w]) D.ab[w] = (const float*)((char*)b->workspace + p->ws_ab[w]);
  }
  for (int a = 0; a < 3; ++a) { D.slo[a] = p->slo[a]; D.shi[a] = p->shi[a]; }
  return D;
}

static int forward_impl(const fdtdx_plan* p, const fdtdx_buffers* b, int32_t t0, int32_t t1, int32_t rec,
                        void* stream, hipEvent_t* ev);

// ---------------------------------------------------------------- forward without allocation
// fdtdx_forward_noalloc runs fdtdx_forward with its workspace (the H_prev snapshot of the detector
// boxes and the detectors' per-block partial sums) in a static device array instead of a caller
// buffer. In host code `scratch` names a host-side stand-in; its device address on the current
// device comes from hipGetSymbolAddress, looked up once per device and kept.
constexpr size_t kScratchBytes = 32u << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static int scratch_address(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(FDTDX_EINVAL, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(FDTDX_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return FDTDX_OK;
}

int fdtdx_forward_noalloc(const fdtdx_plan* p, const fdtdx_buffers* b, int32_t t0, int32_t t1, int32_t rec,
                          void* stream) {
  if (!p || !b) return fail(FDTDX_EINVAL, "null argument");
  if (p->use_fused || p->ec == 9 || p->mc == 9)
    return fail(FDTDX_ENOTSUP, "fdtdx_forward_noalloc: single-pass schedules and full-tensor materials need a "
                               "caller workspace (fdtdx_forward)");
  if (p->ws_fwd_end > kScratchBytes)
    return fail(FDTDX_EINVAL, "fdtdx_forward_noalloc: the forward needs %zu bytes of workspace, above the %zu-byte "
                              "static scratch (use fdtdx_forward with a caller workspace)",
                p->ws_fwd_end, kScratchBytes);
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  fdtdx_buffers b2 = *b;
  b2.workspace = ws;
  return forward_impl(p, &b2, t0, t1, rec, stream, nullptr);
}

int fdtdx_plan_schedule(const fdtdx_plan* p, int32_t* out) {
  if (!p || !out) return fail(FDTDX_EINVAL, "null argument");
  *out = p->use_fused == 2 ? FDTDX_SCHED_MARCH : p->use_fused ? FDTDX_SCHED_FUSED : p->P.vq ? FDTDX_SCHED_QUAD
                                                                                           : FDTDX_SCHED_ONE;
  return FDTDX_OK;
}

int fdtdx_forward(const fdtdx_plan* p, const fdtdx_buffers* b, int32_t t0, int32_t t1, int32_t rec,
                  void* stream) {
  return forward_impl(p, b, t0, t1, rec, stream, nullptr);
}

int fdtdx_forward_timed(const fdtdx_plan* p, const fdtdx_buffers* b, int32_t t0, int32_t t1, int32_t rec,
                        void* stream, double* sweep_ms) {
  if (!sweep_ms) return fail(FDTDX_EINVAL, "sweep_ms is NULL");
  const int n = t1 - t0;
  sweep_ms[0] = sweep_ms[1] = 0.0;
  if (n <= 0) return forward_impl(p, b, t0, t1, rec, stream, nullptr);
  std::vector<hipEvent_t> ev(3 * (size_t)n);
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  int rc = forward_impl(p, b, t0, t1, rec, stream, ev.data());
  double tot[2] = {0.0, 0.0};
  if (rc == FDTDX_OK) {
    if (hipEventSynchronize(ev.back()) != hipSuccess) rc = fail(FDTDX_EHIP, "hipEventSynchronize failed");
    for (int i = 0; rc == FDTDX_OK && i < n; ++i) {
      for (int h = 0; h < 2; ++h) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[3 * i + h], ev[3 * i + h + 1]) != hipSuccess)
          rc = fail(FDTDX_EHIP, "hipEventElapsedTime");
        tot[h] += ms;
      }
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  sweep_ms[0] = tot[0];
  sweep_ms[1] = tot[1];
  return rc;
}

static int forward_impl(const fdtdx_plan* p, const fdtdx_buffers* b, int32_t t0, int32_t t1, int32_t rec,
                        void* stream, hipEvent_t* ev) {
  if (!p) return fail(FDTDX_EINVAL, "null plan");
  int rc = check_buffers(p, b);
  if (rc) return rc;
  if (t0 < 0 || t1 > p->P.T || t0 > t1) return fail(FDTDX_EINVAL, "step range [%d,%d) outside [0,%d)", t0, t1, p->P.T);
  if (rec && !b->record && p->P.rec_per_step) return fail(FDTDX_EINVAL, "recording requested but record is NULL");
  if (rec && p->cplx) return fail(FDTDX_ENOTSUP, "complex-field plans keep no interface record");
  hipStream_t s = (hipStream_t)stream;
  StepDyn D = make_dyn(p, b);
  char* ws = (char*)b->workspace;
  // complex fields: the imaginary part's sweeps (after the real part's, with no sources)
  StepDyn Di = D;
  if (p->cplx) {
    D.cpart = 1;
    Di.cpart = -1;
    Di.E = b->E + 3 * p->P.N;
    Di.H = b->H + 3 * p->P.N;
    Di.psi = b->psi ? b->psi + p->psi_floats : nullptr;
    Di.ade = b->ade_state ? b->ade_state + p->ade_floats : nullptr;
  }
  if (t1 > t0 && (p->P.tsig[0] || p->P.tsig[1])) HIP_TRY(launch_tensor_ab(p->P, D, s));
  float* Eb[2] = {b->E, p->use_fused ? (float*)(ws + p->ws_Ealt) : b->E};
  float* Hb[2] = {b->H, p->use_fused ? (float*)(ws + p->ws_Halt) : b->H};
  float* Pb[2] = {b->psi, p->use_fused ? (float*)(ws + p->ws_psialt) : b->psi};
  for (int n = t0; n < t1; ++n) {
    bool any = false;
    for (int q = 0; q < p->ndet; ++q) any = any || p->tti[q][n] >= 0;
    float* snap = (any && p->has_snap) ? (float*)(ws + p->ws_snapH) : nullptr;
    char* recp = (rec && p->P.rec_per_step && p->rec_slot_of[n] >= 0)
                     ? (char*)b->record + (long long)p->rec_slot_of[n] * p->P.rec_slot_bytes
                     : nullptr;
    const int cur = p->use_fused ? ((n - t0) & 1) : 0;
    const int nxt = p->use_fused ? (cur ^ 1) : 0;
    if (ev) HIP_TRY(hipEventRecord(ev[3 * (n - t0)], s));
    if (p->use_fused == 2) {
      MarchDyn M = p->md;
      M.E_in = Eb[cur]; M.H_in = Hb[cur]; M.E_out = Eb[nxt]; M.H_out = Hb[nxt];
      M.ie = b->inv_eps; M.im = b->inv_mu;
      M.psi_in = Pb[cur]; M.psi_out = Pb[nxt];
      M.ade = b->ade_state; M.record = recp; M.snap = snap;
      for (int a = 0; a < 3; ++a) { M.slo[a] = p->slo[a]; M.shi[a] = p->shi[a]; }
      M.n = n;
      HIP_TRY(launch_march(p->P, M, p->ec, p->mc, s));
    } else if (p->use_fused) {
      FusedDyn F = p->fd;
      F.E_in = Eb[cur]; F.H_in = Hb[cur]; F.E_out = Eb[nxt]; F.H_out = Hb[nxt];
      F.ie = b->inv_eps; F.im = b->inv_mu;
      F.psi_in = Pb[cur]; F.psi_out = Pb[nxt];
      F.ade = b->ade_state; F.record = recp; F.snap = snap;
      for (int a = 0; a < 3; ++a) { F.slo[a] = p->slo[a]; F.shi[a] = p->shi[a]; }
      F.n = n;
      HIP_TRY(launch_fused(p->P, F, p->ec, p->mc, p->fused_lds, s));
    } else {
      D.n = n;
      D.snap = snap;
      D.record = p->P.rec_narrow ? nullptr : recp;  // narrow layers are encoded by k_record
      Di.n = n;
      Di.snap = snap ? (float*)(ws + p->ws_snapE) : nullptr;  // H_prev of the imaginary part
      if (p->P.vq) HIP_TRY(launch_update_E4(p->P, D, p->ec, s));
      else HIP_TRY(launch_update_E(p->P, D, p->ec, s));
      if (p->cplx) HIP_TRY(launch_update_E(p->P, Di, p->ec, s));
      if (ev) HIP_TRY(hipEventRecord(ev[3 * (n - t0) + 1], s));
      if (p->P.vq) HIP_TRY(launch_update_H4(p->P, D, p->mc, s));
      else HIP_TRY(launch_update_H(p->P, D, p->mc, s));
      if (p->cplx) HIP_TRY(launch_update_H(p->P, Di, p->mc, s));
      if (p->P.rec_narrow && recp) {
        D.record = recp;
        HIP_TRY(launch_record(p->P, D, p->iface_cells, s));
      }
    }
    if (ev) {
      if (p->use_fused) HIP_TRY(hipEventRecord(ev[3 * (n - t0) + 1], s));
      HIP_TRY(hipEventRecord(ev[3 * (n - t0) + 2], s));
    }
    // the detectors of step n in grouped launches of up to kDetGroup (a chain of accumulating
    // detectors stays in one group); one finalize launch per group for the reduced values
    DetGroup G;
    G.count = G.nfin = 0;
    G.start[0] = 0;
    auto flush = [&]() -> int {
      const hipError_t e = launch_detector_group(p->P, p->det_dev, G, s);
      G.count = G.nfin = 0;
      return e == hipSuccess ? FDTDX_OK : fail(FDTDX_EHIP, "detector launch: %s", hipGetErrorString(e));
    };
    for (int q = 0; q < p->ndet; ++q) {
      const int slot = p->tti[q][n];
      if (slot < 0 || p->det[q].inv) continue;
      if (G.count == kDetGroup) {  // full: move a trailing accumulation chain to the next group
        int h = G.count;
        if (p->det[q].accum) {
          while (h > 0 && p->det[G.q[h - 1]].accum) --h;
          --h;  // the chain head
        }
        if (h <= 0) return fail(FDTDX_ENOTSUP, "more than %d accumulating detectors in one chain", kDetGroup);
        const int tail = G.count - h;
        DetGroup R;
        R.count = 0; R.nfin = 0; R.start[0] = 0;
        for (int t = 0; t < tail; ++t) {
          const int a = h + t;
          R.q[t] = G.q[a]; R.slot[t] = G.slot[a]; R.dd[t] = G.dd[a];
          R.start[t + 1] = R.start[t] + (G.start[a + 1] - G.start[a]);
        }
        R.count = tail;
        G.count = h;
        int nf = 0;
        for (int b2 = 0; b2 < G.nfin; ++b2) if (G.fin[b2] < h) G.fin[nf++] = G.fin[b2];
        G.nfin = nf;
        if ((rc = flush())) return rc;
        G = R;
        for (int a = 0; a < G.count; ++a)
          if (reduced_values(p->det[G.q[a]]) && !p->det[G.q[a]].accum) G.fin[G.nfin++] = a;
      }
      const int a = G.count;
      DetDyn& dd = G.dd[a];
      memset(&dd, 0, sizeof(dd));
      dd.E = Eb[nxt]; dd.H = Hb[nxt]; dd.Hprev = (const float*)(ws + p->ws_snapH);
      if (p->cplx) {
        dd.Ei = Di.E; dd.Hi = Di.H; dd.Hprevi = (const float*)(ws + p->ws_snapE);
      }
      for (int c = 0; c < 3; ++c) { dd.slo[c] = p->slo[c]; dd.shi[c] = p->shi[c]; }
      dd.ie = b->inv_eps; dd.im = b->inv_mu; dd.ec = p->ec; dd.mc = p->mc;
      dd.state = b->detector_state[q];
      dd.slot = p->det[q].kind == FDTDX_DET_PHASOR ? 0 : slot;
      dd.n = n;
      dd.partial = (double*)(ws + p->ws_partial) + (size_t)q * kDetMaxValues * kDetMaxBlocks;
      G.q[a] = q;
      G.slot[a] = dd.slot;
      G.start[a + 1] = G.start[a] + detector_blocks(p->det[q]);
      if (reduced_values(p->det[q]) && !p->det[q].accum) G.fin[G.nfin++] = a;
      G.count = a + 1;
    }
    if (G.count > 0 && (rc = flush())) return rc;
  }
  if (p->use_fused && ((t1 - t0) & 1)) {
    HIP_TRY(hipMemcpyAsync(b->E, Eb[1], 3 * (size_t)p->P.N * 4, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->H, Hb[1], 3 * (size_t)p->P.N * 4, hipMemcpyDeviceToDevice, s));
    if (p->psi_floats) HIP_TRY(hipMemcpyAsync(b->psi, Pb[1], p->psi_floats * 4, hipMemcpyDeviceToDevice, s));
  }
  return FDTDX_OK;
}

int fdtdx_reverse(const fdtdx_plan* p, const fdtdx_buffers* b, int32_t t1, int32_t t0, int32_t flags, void* stream) {
  if (!p) return fail(FDTDX_EINVAL, "null plan");
  int rc = check_buffers(p, b);
  if (rc) return rc;
  if (t0 < 0 || t1 > p->P.T || t0 > t1) return fail(FDTDX_EINVAL, "step range [%d,%d) outside [0,%d)", t0, t1, p->P.T);
  if (p->P.nade) return fail(FDTDX_ENOTSUP, "reverse pass with dispersive (ADE) media is not supported (fdtd.py:98-102)");
  if (p->cplx) return fail(FDTDX_ENOTSUP, "the reverse run of complex-field plans is not supported");
  if (p->P.rec_per_step && !b->record) return fail(FDTDX_EINVAL, "record is NULL");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)b->workspace;
  StepDyn D = make_dyn(p, b);
  const bool reset = flags & 1, record = flags & 2;
  if (t1 > t0 && (p->P.tsig[0] || p->P.tsig[1])) HIP_TRY(launch_tensor_ab(p->P, D, s));
  for (int n = t1 - 1; n >= t0; --n) {
    D.n = n;
    D.record = p->P.rec_per_step ? (char*)b->record + (long long)p->rec_arr[n] * p->P.rec_slot_bytes : nullptr;
    D.rec_interp = p->P.rec_per_step && p->rec_slot_of[n] < 0;
    D.rec_b = D.rec_interp ? D.record + p->P.rec_slot_bytes : nullptr;
    D.rec_f = p->rec_factor[n];
    D.snap = nullptr;
    HIP_TRY(launch_restore(p->P, D, p->iface_cells, s));
    bool any = false;
    for (int q = 0; q < p->ndet; ++q) any = any || (record && p->det[q].inv && p->tti[q][n] >= 0);
    if (any)  // the H before the undo, for the inverse detectors' time-centred H
      HIP_TRY(launch_snapshot(p->P, b->H, (float*)(ws + p->ws_snapH), p->slo, p->shi, s));
    HIP_TRY(launch_reverse(p->P, D, p->ec, p->mc, s));
    if (reset) {
      for (int sl = 0; sl < p->P.nslab; ++sl) {
        const SlabK& S = p->P.slab[sl];
        int lo[3] = {0, 0, 0}, hi[3] = {p->P.nx, p->P.ny, p->P.nz};
        lo[S.axis] = S.lo;
        hi[S.axis] = S.hi;
        HIP_TRY(launch_zero_box(p->P, b->E, lo, hi, s));
        HIP_TRY(launch_zero_box(p->P, b->H, lo, hi, s));
      }
    }
    if (!any) continue;
    for (int q = 0; q < p->ndet; ++q) {
      const int slot = p->tti[q][n];
      if (slot < 0 || !p->det[q].inv) continue;
      DetDyn dd;
      memset(&dd, 0, sizeof(dd));
      dd.E = b->E; dd.H = b->H; dd.Hprev = (const float*)(ws + p->ws_snapH);
      for (int a = 0; a < 3; ++a) { dd.slo[a] = p->slo[a]; dd.shi[a] = p->shi[a]; }
      dd.ie = b->inv_eps; dd.im = b->inv_mu; dd.ec = p->ec; dd.mc = p->mc;
      dd.state = b->detector_state[q];
      dd.slot = p->det[q].kind == FDTDX_DET_PHASOR ? 0 : slot;
      dd.n = n;
      dd.partial = (double*)(ws + p->ws_partial) + (size_t)q * kDetMaxValues * kDetMaxBlocks;
      HIP_TRY(launch_detector(p->P, p->det[q], dd, detector_blocks(p->det[q]), s));
    }
  }
  return FDTDX_OK;
}

int fdtdx_reverse_adjoint(const fdtdx_plan* p, const fdtdx_buffers* b, const fdtdx_adj_buffers* adj, int32_t t1,
                          int32_t t0, void* stream) {
  if (!p) return fail(FDTDX_EINVAL, "null plan");
  int rc = check_buffers(p, b);
  if (rc) return rc;
  if (!adj || !adj->lam_E || !adj->lam_H || !adj->grad_inv_eps)
    return fail(FDTDX_EINVAL, "lam_E, lam_H and grad_inv_eps are required");
  if (p->cplx) return fail(FDTDX_ENOTSUP, "gradients of complex-field plans are not supported");
  if ((p->ec == 9 || p->mc == 9) && (p->P.loss[0].comps || p->P.loss[1].comps))
    return fail(FDTDX_ENOTSUP, "gradients in full-tensor media with conductivity are not supported");
  if (p->psi_floats && !adj->lam_psi) return fail(FDTDX_EINVAL, "lam_psi is NULL");
  if (t0 < 0 || t1 > p->P.T || t0 > t1) return fail(FDTDX_EINVAL, "step range [%d,%d) outside [0,%d)", t0, t1, p->P.T);
  if (p->P.nade) return fail(FDTDX_ENOTSUP, "reverse pass with dispersive (ADE) media is not supported (fdtd.py:98-102)");
  if (p->P.rec_per_step && !b->record) return fail(FDTDX_EINVAL, "record is NULL");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)b->workspace;
  StepDyn D = make_dyn(p, b);
  float* lpsi_a = adj->lam_psi;
  float* lpsi_b = (float*)(ws + p->ws_lpsi);
  for (int n = t1 - 1; n >= t0; --n) {
    D.n = n;
    D.record = p->P.rec_per_step ? (char*)b->record + (long long)p->rec_arr[n] * p->P.rec_slot_bytes : nullptr;
    D.rec_interp = p->P.rec_per_step && p->rec_slot_of[n] < 0;
    D.rec_b = D.rec_interp ? D.record + p->P.rec_slot_bytes : nullptr;
    D.rec_f = p->rec_factor[n];
    D.snap = nullptr;
    HIP_TRY(launch_restore(p->P, D, p->iface_cells, s));
    bool any = false;
    for (int q = 0; q < p->ndet; ++q) {
      if (!adj->detector_cotangent || !adj->detector_cotangent[q] || p->det[q].inv) continue;
      if (p->tti[q][n] < 0) continue;
      any = true;
    }
    if (any) {
      HIP_TRY(launch_snapshot(p->P, b->E, (float*)(ws + p->ws_snapE), p->slo, p->shi, s, b->H,
                              (float*)(ws + p->ws_snapH)));
    }
    // diagonal materials: only the H half-step here; the E half-step runs inside pass A of the
    // adjoint (it reads the same H^n and psi), after the detector cotangents, which read H^n and
    // the E^{n+1} snapshot but not E
    const bool fuse = p->ec != 9 && p->mc != 9;
    if (fuse) HIP_TRY(launch_reverse_H(p->P, D, p->mc, s));
    else HIP_TRY(launch_reverse(p->P, D, p->ec, p->mc, s));
    DetAdjDyn dd[FDTDX_MAX_DETECTORS];
    bool on[FDTDX_MAX_DETECTORS];
    for (int q = 0; q < p->ndet; ++q) {
      on[q] = any && adj->detector_cotangent && adj->detector_cotangent[q] && !p->det[q].inv && p->tti[q][n] >= 0;
      if (!on[q]) continue;
      DetAdjDyn& x = dd[q];
      memset(&x, 0, sizeof(x));
      x.E1 = (const float*)(ws + p->ws_snapE);
      x.H1 = (const float*)(ws + p->ws_snapH);
      x.H0 = b->H;
      for (int a = 0; a < 3; ++a) {
        x.slo[a] = p->slo[a]; x.shi[a] = p->shi[a];
        x.elo[a] = p->elo[q][a]; x.ehi[a] = p->ehi[q][a];
      }
      x.ie = b->inv_eps; x.ec = p->ec; x.im = b->inv_mu; x.mc = p->mc; x.grad = adj->grad_inv_eps;
      x.cot = adj->detector_cotangent[q];
      x.slot = p->det[q].kind == FDTDX_DET_PHASOR ? 0 : p->tti[q][n];
      x.n = n;
      x.g = (float*)(ws + p->g_off[q]);
      x.half_vH = (float*)(ws + p->half_off[q]);
      x.lamE = adj->lam_E; x.lamH = adj->lam_H;
      HIP_TRY(launch_detector_adjoint(p->P, p->det[q], x, s));
    }
    AdjDyn A;
    A.lamE = adj->lam_E; A.lamH = adj->lam_H;
    A.lpsi_in = lpsi_a; A.lpsi_out = lpsi_b;
    A.grad = adj->grad_inv_eps;
    A.H = b->H; A.E = b->E; A.psi = b->psi; A.ie = b->inv_eps; A.im = b->inv_mu;
    A.K = (p->ec == 9 || p->mc == 9) ? (float*)(ws + p->ws_K) : nullptr;
    A.W = (p->ec == 9 || p->mc == 9) ? (float*)(ws + p->ws_W) : nullptr;
    A.Erev = fuse ? b->E : nullptr;
    A.n = n;
    HIP_TRY(launch_adjoint(p->P, A, p->ec, p->mc, s));
    std::swap(lpsi_a, lpsi_b);
    for (int q = 0; q < p->ndet; ++q)
      if (on[q]) HIP_TRY(launch_add_half_vH(p->P, dd[q], s));
  }
  if (lpsi_a != adj->lam_psi && p->psi_floats)
    HIP_TRY(hipMemcpyAsync(adj->lam_psi, lpsi_a, p->psi_floats * 4, hipMemcpyDeviceToDevice, s));
  return FDTDX_OK;
}

// ---------------------------------------------------------------- multi-instance batch
struct fdtdx_batch {
  int ec = 1, mc = 0;
  int grid = 0, det_grid = 1024;
  BatchTable* dtab = nullptr;      // device table
  int* dbase = nullptr;            // device step counters, one per slot
  int* dflags = nullptr;           // device health flags, one per slot
  int* hflags = nullptr;           // pinned host copy of the flags, refreshed after every chunk
  float blowup = 0.f;
  BatchTable* stage[2] = {nullptr, nullptr};  // pinned staging for table uploads
  hipEvent_t staged[2] = {nullptr, nullptr};
  int next_stage = 0;
  bool dirty = true;
  struct Slot {
    const fdtdx_plan* plan = nullptr;
    fdtdx_buffers buf;
    void* det_state[FDTDX_MAX_DETECTORS];
    int t0 = 0, t1 = 0, rec = 0;
    int n = 0;                     // step index after the enqueued chunks
    bool pending_set = false;      // counter must be initialised before the next chunk
  };
  Slot slot[kMaxBatch];
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;  // (steps, with detectors, grid) -> executable

  ~fdtdx_batch() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    for (int i = 0; i < 2; ++i) {
      if (staged[i]) (void)hipEventDestroy(staged[i]);
      if (stage[i]) (void)hipHostFree(stage[i]);
    }
    if (dtab) (void)hipFree(dtab);
    if (dbase) (void)hipFree(dbase);
    if (dflags) (void)hipFree(dflags);
    if (hflags) (void)hipHostFree(hflags);
  }
};

int fdtdx_batch_create(int32_t eps_comps, int32_t mu_comps, float blowup_limit, fdtdx_batch** out) {
  std::lock_guard<std::mutex> lk(g_device_mu);
  if (!out) return fail(FDTDX_EINVAL, "null argument");
  *out = nullptr;
  if (!(blowup_limit >= 0.f)) return fail(FDTDX_EINVAL, "blowup_limit must be >= 0");
  if (eps_comps != 1 && eps_comps != 3) return fail(FDTDX_ENOTSUP, "eps_comps=%d", eps_comps);
  if (mu_comps != 0 && mu_comps != 1 && mu_comps != 3) return fail(FDTDX_ENOTSUP, "mu_comps=%d", mu_comps);
  fdtdx_batch* b = new fdtdx_batch();
  b->ec = eps_comps;
  b->mc = mu_comps;
  b->blowup = blowup_limit;
  auto bail = [&](int code) { delete b; return code; };
  hipError_t e = hipMalloc((void**)&b->dtab, sizeof(BatchTable));
  if (e == hipSuccess) e = hipMalloc((void**)&b->dbase, sizeof(int) * kMaxBatch);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dflags, sizeof(int) * kMaxBatch);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->hflags, sizeof(int) * kMaxBatch, hipHostMallocDefault);
  if (e == hipSuccess) memset(b->hflags, 0, sizeof(int) * kMaxBatch);
  // Counters and flags start defined; later changes go through kernels only (graph-safe).
  if (e == hipSuccess) e = hipMemset(b->dbase, 0, sizeof(int) * kMaxBatch);
  if (e == hipSuccess) e = hipMemset(b->dflags, 0, sizeof(int) * kMaxBatch);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  for (int i = 0; e == hipSuccess && i < 2; ++i) {
    e = hipHostMalloc((void**)&b->stage[i], sizeof(BatchTable), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) return bail(fail(FDTDX_EHIP, "batch allocation: %s", hipGetErrorString(e)));
  b->grid = batch_sweep_grid(0);  // sized to the batch at each upload
  *out = b;
  return FDTDX_OK;
}

void fdtdx_batch_destroy(fdtdx_batch* batch) {
  std::lock_guard<std::mutex> lk(g_device_mu);
  delete batch;
}

int fdtdx_batch_attach(fdtdx_batch* b, int32_t slot, const fdtdx_plan* p, const fdtdx_buffers* buf, int32_t t0,
                       int32_t t1, int32_t rec) {
  if (!b || !p || !buf) return fail(FDTDX_EINVAL, "null argument");
  if (slot < 0 || slot >= kMaxBatch) return fail(FDTDX_EINVAL, "slot %d outside [0,%d)", slot, kMaxBatch);
  if (p->ec != b->ec || p->mc != b->mc)
    return fail(FDTDX_EINVAL, "plan material class (%d,%d) differs from the batch's (%d,%d)", p->ec, p->mc, b->ec, b->mc);
  if (p->cplx) return fail(FDTDX_ENOTSUP, "complex-field plans do not run in batches");
  int rc = check_buffers(p, buf);
  if (rc) return rc;
  if (t0 < 0 || t1 > p->P.T || t0 > t1) return fail(FDTDX_EINVAL, "step range [%d,%d) outside [0,%d)", t0, t1, p->P.T);
  if (rec && !buf->record && p->P.rec_per_step) return fail(FDTDX_EINVAL, "recording requested but record is NULL");
  if (rec && p->P.rec_narrow)
    return fail(FDTDX_ENOTSUP, "a batch records float32 interface layers only (the plan stores narrow formats)");
  for (int q = 0; q < p->ndet; ++q)
    if (p->det[q].inv) return fail(FDTDX_ENOTSUP, "detector %d: inverse detectors are recorded by fdtdx_reverse only", q);
  fdtdx_batch::Slot& s = b->slot[slot];
  s.plan = p;
  s.buf = *buf;
  for (int q = 0; q < p->ndet; ++q) s.det_state[q] = buf->detector_state[q];
  s.buf.detector_state = s.det_state;
  s.t0 = t0; s.t1 = t1; s.rec = rec;
  s.n = t0;
  s.pending_set = true;
  b->dirty = true;
  return FDTDX_OK;
}

int fdtdx_batch_detach(fdtdx_batch* b, int32_t slot) {
  if (!b) return fail(FDTDX_EINVAL, "null batch");
  if (slot < 0 || slot >= kMaxBatch) return fail(FDTDX_EINVAL, "slot %d outside [0,%d)", slot, kMaxBatch);
  b->slot[slot].plan = nullptr;
  b->slot[slot].pending_set = false;
  b->dirty = true;
  return FDTDX_OK;
}

int fdtdx_batch_step_index(const fdtdx_batch* b, int32_t slot, int32_t* n) {
  if (!b || !n) return fail(FDTDX_EINVAL, "null argument");
  if (slot < 0 || slot >= kMaxBatch) return fail(FDTDX_EINVAL, "slot %d outside [0,%d)", slot, kMaxBatch);
  if (!b->slot[slot].plan) return fail(FDTDX_EINVAL, "slot %d is not attached", slot);
  *n = b->slot[slot].n;
  return FDTDX_OK;
}

// Fill the host table from the attached slots (slot order) and enqueue its upload.
static int batch_upload(fdtdx_batch* b, hipStream_t s) {
  const int k = b->next_stage;
  b->next_stage ^= 1;
  HIP_TRY(hipEventSynchronize(b->staged[k]));  // the previous upload from this buffer has finished
  BatchTable* T = b->stage[k];
  int e = 0;
  T->tile_start[0] = 0;
  T->det_start[0] = 0;
  T->fin_start[0] = 0;
  for (int sl = 0; sl < kMaxBatch; ++sl) {
    const fdtdx_batch::Slot& S = b->slot[sl];
    if (!S.plan) continue;
    const fdtdx_plan* p = S.plan;
    BatchInst& I = T->inst[e];
    memset(&I, 0, sizeof(I));
    I.P = p->P;
    I.D = make_dyn(p, &S.buf);
    char* ws = (char*)S.buf.workspace;
    I.D.snap = p->has_snap ? (float*)(ws + p->ws_snapH) : nullptr;
    I.rec_base = (S.rec && p->P.rec_per_step) ? (char*)S.buf.record : nullptr;
    I.stop = S.t1;
    I.slot = sl;
    I.ec = p->ec; I.mc = p->mc;
    I.ndet = p->ndet;
    int nb = 0, nf = 0;
    for (int q = 0; q < p->ndet; ++q) {
      I.det[q] = p->det[q];
      I.det_nb[q] = detector_blocks(p->det[q]);
      nb += I.det_nb[q];
      if (reduced_values(p->det[q]) && !p->det[q].accum) ++nf;
      DetDyn& dd = I.dd[q];
      dd.E = S.buf.E; dd.H = S.buf.H; dd.Hprev = (const float*)(ws + p->ws_snapH);
      for (int a = 0; a < 3; ++a) { dd.slo[a] = p->slo[a]; dd.shi[a] = p->shi[a]; }
      dd.ie = S.buf.inv_eps; dd.im = S.buf.inv_mu; dd.ec = p->ec; dd.mc = p->mc;
      dd.state = S.det_state[q];
      dd.partial = (double*)(ws + p->ws_partial) + (size_t)q * kDetMaxValues * kDetMaxBlocks;
    }
    T->tile_start[e + 1] = T->tile_start[e] + p->P.nblocks;
    T->det_start[e + 1] = T->det_start[e] + nb;
    T->fin_start[e + 1] = T->fin_start[e] + nf;
    ++e;
  }
  T->ninst = e;
  b->grid = batch_sweep_grid(T->tile_start[e]);
  T->blowup = b->blowup;
  T->flags = b->dflags;
  const size_t bytes = offsetof(BatchTable, inst) + (size_t)e * sizeof(BatchInst);
  HIP_TRY(hipMemcpyAsync(b->dtab, T, bytes, hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(b->staged[k], s));
  b->dirty = false;
  return FDTDX_OK;
}

static int batch_chunk(fdtdx_batch* b, int steps, bool det, hipStream_t s) {
  for (int r = 0; r < steps; ++r) {
    HIP_TRY(launch_batch_sweeps(b->dtab, b->dbase, r, b->ec, b->mc, b->grid, s));
    if (det) HIP_TRY(launch_batch_detectors(b->dtab, b->dbase, r, b->det_grid, true, s));
  }
  HIP_TRY(launch_batch_advance(b->dtab, b->dbase, steps, s));
  return FDTDX_OK;
}

int fdtdx_batch_advance(fdtdx_batch* b, int32_t steps, int32_t use_graph, void* stream) {
  if (!b) return fail(FDTDX_EINVAL, "null batch");
  if (steps <= 0) return fail(FDTDX_EINVAL, "steps must be positive");
  hipStream_t s = (hipStream_t)stream;
  if (use_graph && !s) return fail(FDTDX_EINVAL, "graph capture needs a created stream, not the default stream");
  int rc;
  if (b->dirty && (rc = batch_upload(b, s))) return rc;
  BatchSet set;
  memset(&set, 0, sizeof(set));
  bool det = false, any = false;
  for (int sl = 0; sl < kMaxBatch; ++sl) {
    fdtdx_batch::Slot& S = b->slot[sl];
    if (!S.plan) continue;
    any = true;
    det = det || S.plan->ndet > 0;
    if (S.pending_set) {
      set.slot[set.count] = sl;
      set.value[set.count] = S.t0;
      ++set.count;
      S.pending_set = false;
    }
  }
  if (!any) return FDTDX_OK;
  HIP_TRY(launch_batch_set(b->dbase, b->dflags, set, s));
  if (use_graph) {
    const auto key = std::make_tuple((int)steps, det ? 1 : 0, b->grid);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
      hipGraph_t g = nullptr;
      std::unique_lock<std::mutex> lk(g_device_mu);
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      rc = batch_chunk(b, steps, det, s);
      hipError_t e = hipStreamEndCapture(s, &g);
      lk.unlock();
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (e != hipSuccess) return fail(FDTDX_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      hipGraphExec_t x = nullptr;
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(FDTDX_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      it = b->graphs.emplace(key, x).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, s));
  } else if ((rc = batch_chunk(b, steps, det, s))) {
    return rc;
  }
  HIP_TRY(hipMemcpyAsync(b->hflags, b->dflags, sizeof(int) * kMaxBatch, hipMemcpyDeviceToHost, s));
  for (int sl = 0; sl < kMaxBatch; ++sl) {
    fdtdx_batch::Slot& S = b->slot[sl];
    if (S.plan && S.n < S.t1) S.n = S.n + steps < S.t1 ? S.n + steps : S.t1;
  }
  return FDTDX_OK;
}

int fdtdx_batch_flags(const fdtdx_batch* b, int32_t* flags) {
  if (!b || !flags) return fail(FDTDX_EINVAL, "null argument");
  for (int sl = 0; sl < kMaxBatch; ++sl) flags[sl] = b->slot[sl].pending_set ? 0 : b->hflags[sl];
  return FDTDX_OK;
}

}  // extern "C"
