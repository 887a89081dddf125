// Host-side AddressSanitizer run of the plan builder and the entry points' argument checks.
//
// Built by fdtdx_amd/build.py (build_plan_asan) with plan.hip's host code instrumented
// (-Xarch_host -fsanitize=address; the device code is not), and run on a GPU by
// tests/test_gpu_plan_asan.py. It creates a valid plan (CPML on every face, a dipole, flux, phasor
// and energy detectors, the interface record), runs a forward with recording and the reverse
// adjoint, then hands fdtdx_plan_create one malformed descriptor after another: each must be
// rejected with a status and a message, without a memory error in the builder's partial-build
// clean-up. The last line printed is "plan_asan: <n> cases passed" when every case behaved.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "fdtdx_hip.h"

namespace {

constexpr int kN = 16, kT = 24, kPml = 4, kFreqs = 2;

int g_failed = 0, g_cases = 0;

void expect(bool ok, const std::string& what) {
  ++g_cases;
  if (!ok) {
    ++g_failed;
    std::printf("FAILED: %s (last error: %s)\n", what.c_str(), fdtdx_last_error());
  }
  std::fflush(stdout);
}

// Host tables of the valid scene; descriptors point into them.
struct Scene {
  std::vector<float> coef[6][6];  // [slab][b_E, a_E, ikm1_E, b_H, a_H, ikm1_H]
  std::vector<float> src_table;
  std::vector<int32_t> tti;
  std::vector<float> cos_t, sin_t, window;
  fdtdx_cpml_desc cpml[6];
  fdtdx_source_desc src[1];
  fdtdx_detector_desc det[3];

  fdtdx_plan_desc desc() {
    fdtdx_plan_desc d;
    std::memset(&d, 0, sizeof(d));
    d.nx = d.ny = d.nz = kN;
    d.num_steps = kT;
    d.courant = 0.99f / std::sqrt(3.0f);
    d.eps_comps = 1;
    d.mu_comps = 0;
    for (int s = 0; s < 6; ++s) {
      const int axis = s / 2, side = s % 2;
      for (int t = 0; t < 6; ++t) {
        coef[s][t].assign(kPml, 0.f);
        for (int q = 0; q < kPml; ++q) {
          const float depth = (side ? q + 1.f : kPml - q) / kPml;  // 1 at the outer face
          coef[s][t][q] = t % 3 == 0 ? std::exp(-0.3f * depth) : (t % 3 == 1 ? -0.2f * depth : 0.f);
        }
      }
      fdtdx_cpml_desc& c = cpml[s];
      std::memset(&c, 0, sizeof(c));
      c.axis = axis;
      c.side = side;
      c.lo = side ? kN - kPml : 0;
      c.hi = side ? kN : kPml;
      c.kappa_is_one = 1;
      c.b_E = coef[s][0].data(); c.a_E = coef[s][1].data(); c.ikm1_E = coef[s][2].data();
      c.b_H = coef[s][3].data(); c.a_H = coef[s][4].data(); c.ikm1_H = coef[s][5].data();
    }
    d.num_cpml = 6;
    d.cpml = cpml;
    src_table.resize(kT);
    for (int n = 0; n < kT; ++n) src_table[n] = -0.1f * std::exp(-0.5f * (n - 8.f) * (n - 8.f) / 9.f);
    std::memset(src, 0, sizeof(src));
    src[0].kind = FDTDX_SRC_POINT;
    src[0].field = 0;
    src[0].component = 2;
    for (int a = 0; a < 3; ++a) { src[0].lo[a] = kN / 2 - 2; src[0].hi[a] = src[0].lo[a] + 1; }
    src[0].table = src_table.data();
    d.num_sources = 1;
    d.sources = src;
    tti.resize(kT);
    for (int n = 0; n < kT; ++n) tti[n] = n;
    cos_t.resize(kFreqs * kT);
    sin_t.resize(kFreqs * kT);
    window.assign(kT, 1.f);
    for (int f = 0; f < kFreqs; ++f)
      for (int n = 0; n < kT; ++n) {
        cos_t[f * kT + n] = std::cos(0.3f * (f + 1) * n);
        sin_t[f * kT + n] = std::sin(0.3f * (f + 1) * n);
      }
    std::memset(det, 0, sizeof(det));
    // flux plane (reduced), phasor plane (reduced, Ez and Hy), energy box (reduced)
    det[0].kind = FDTDX_DET_FLUX;
    det[1].kind = FDTDX_DET_PHASOR;
    det[2].kind = FDTDX_DET_ENERGY;
    for (int q = 0; q < 3; ++q) {
      for (int a = 0; a < 3; ++a) { det[q].lo[a] = kPml + 1; det[q].hi[a] = kN - kPml - 1; }
      det[q].reduce_volume = 1;
      det[q].exact_interpolation = 1;
      det[q].weight = 1.f;
      det[q].sign = 1.f;
      det[q].time_to_idx = tti.data();
    }
    det[0].lo[0] = kN - kPml - 2; det[0].hi[0] = det[0].lo[0] + 1;
    det[0].axis = 0;
    det[1].lo[0] = kN - kPml - 3; det[1].hi[0] = det[1].lo[0] + 1;
    det[1].num_components = 2;
    det[1].components[0] = 2;
    det[1].components[1] = 4;
    det[1].num_freqs = kFreqs;
    det[1].scale = 1.f;
    det[1].cos_table = cos_t.data();
    det[1].sin_table = sin_t.data();
    det[1].window = window.data();
    d.num_detectors = 3;
    d.detectors = det;
    return d;
  }
};

// Device buffers of one solve of the valid scene.
struct Solve {
  std::vector<void*> allocs;
  fdtdx_buffers b;
  fdtdx_adj_buffers adj;
  void* det_state[3] = {nullptr, nullptr, nullptr};
  const void* det_cot[3] = {nullptr, nullptr, nullptr};
  size_t N = (size_t)kN * kN * kN;

  float* dev(size_t floats, float fill) {
    std::vector<float> h(floats ? floats : 1, fill);
    void* p = nullptr;
    if (hipMalloc(&p, h.size() * 4) != hipSuccess) return nullptr;
    hipMemcpy(p, h.data(), h.size() * 4, hipMemcpyHostToDevice);
    allocs.push_back(p);
    return (float*)p;
  }
  bool make(const fdtdx_plan* pl) {
    size_t psi = 0, ws = 0, slot = 0, slots = 0;
    if (fdtdx_psi_floats(pl, &psi) || fdtdx_workspace_bytes(pl, &ws) || fdtdx_record_slot_bytes(pl, &slot) ||
        fdtdx_record_slots(pl, &slots))
      return false;
    std::memset(&b, 0, sizeof(b));
    b.E = dev(3 * N, 0.f);
    b.H = dev(3 * N, 0.f);
    b.inv_eps = dev(N, 0.5f);
    b.psi = dev(psi, 0.f);
    det_state[0] = dev(kT, 0.f);
    det_state[1] = dev(2 * kFreqs * 2, 0.f);
    det_state[2] = dev(kT, 0.f);
    b.detector_state = det_state;
    b.record = dev(slot * slots / 4 + 1, 0.f);
    b.workspace = dev(ws / 4 + 1, 0.f);
    std::memset(&adj, 0, sizeof(adj));
    adj.lam_E = dev(3 * N, 0.f);
    adj.lam_H = dev(3 * N, 0.f);
    adj.lam_psi = dev(psi, 0.f);
    adj.grad_inv_eps = dev(N, 0.f);
    det_cot[0] = dev(kT, 1.f);
    adj.detector_cotangent = det_cot;
    for (void* p : allocs)
      if (!p) return false;
    return true;
  }
  ~Solve() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

void valid_run(Scene& sc) {
  fdtdx_plan_desc d = sc.desc();
  fdtdx_plan* pl = nullptr;
  expect(fdtdx_plan_create(&d, &pl) == FDTDX_OK && pl, "valid plan");
  if (!pl) return;
  {
    Solve s;
    expect(s.make(pl), "buffers of the valid plan");
    expect(fdtdx_forward(pl, &s.b, 0, kT, 1, nullptr) == FDTDX_OK, "forward with recording");
    expect(fdtdx_reverse_adjoint(pl, &s.b, &s.adj, kT, 0, nullptr) == FDTDX_OK, "reverse adjoint");
    expect(hipDeviceSynchronize() == hipSuccess, "device synchronize");
    std::vector<float> g(s.N), flux(kT);
    hipMemcpy(g.data(), s.adj.grad_inv_eps, s.N * 4, hipMemcpyDeviceToHost);
    hipMemcpy(flux.data(), s.det_state[0], kT * 4, hipMemcpyDeviceToHost);
    bool finite = true, nonzero = false;
    for (float v : g) { finite = finite && std::isfinite(v); nonzero = nonzero || v != 0.f; }
    expect(finite && nonzero, "finite, nonzero gradient");
    bool fl = false;
    for (float v : flux) fl = fl || v != 0.f;
    expect(fl, "nonzero flux trace");
    // argument checks of the entry points
    fdtdx_buffers nb = s.b;
    nb.E = nullptr;
    expect(fdtdx_forward(pl, &nb, 0, 1, 0, nullptr) == FDTDX_EINVAL, "forward without E");
    nb = s.b;
    nb.detector_state = nullptr;
    expect(fdtdx_forward(pl, &nb, 0, 1, 0, nullptr) == FDTDX_EINVAL, "forward without detector states");
    expect(fdtdx_forward(pl, &s.b, 0, kT + 1, 0, nullptr) != FDTDX_OK, "forward past the last step");
    expect(fdtdx_reverse_adjoint(pl, &s.b, &s.adj, kT + 1, 0, nullptr) == FDTDX_EINVAL, "adjoint past the last step");
    fdtdx_adj_buffers na = s.adj;
    na.lam_psi = nullptr;
    expect(fdtdx_reverse_adjoint(pl, &s.b, &na, kT, 0, nullptr) == FDTDX_EINVAL, "adjoint without lam_psi");
    int32_t slot = 0;
    expect(fdtdx_record_slot_of(pl, kT, &slot) != FDTDX_OK, "record slot of a step past the end");
  }
  fdtdx_plan_destroy(pl);
}

void malformed(Scene& sc, const char* what, const std::function<void(fdtdx_plan_desc&)>& edit) {
  fdtdx_plan_desc d = sc.desc();
  edit(d);
  fdtdx_plan* pl = nullptr;
  const int rc = fdtdx_plan_create(&d, &pl);
  const bool ok = rc != FDTDX_OK && pl == nullptr && std::strlen(fdtdx_last_error()) > 0;
  expect(ok, std::string("rejects ") + what);
  if (pl) fdtdx_plan_destroy(pl);
}

}  // namespace

int main() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::printf("plan_asan: no GPU\n");
    return 2;
  }
  Scene sc;
  valid_run(sc);
  // every descriptor array below is a copy of the scene's, edited after desc() filled it
  static fdtdx_cpml_desc cp[6];
  static fdtdx_source_desc sr[2];
  static fdtdx_detector_desc dt[3];
  auto cpml_edit = [&](fdtdx_plan_desc& d, const std::function<void(fdtdx_cpml_desc*)>& f) {
    std::memcpy(cp, d.cpml, sizeof(cp));
    f(cp);
    d.cpml = cp;
  };
  auto src_edit = [&](fdtdx_plan_desc& d, const std::function<void(fdtdx_source_desc&)>& f) {
    sr[0] = d.sources[0];
    f(sr[0]);
    d.sources = sr;
  };
  auto det_edit = [&](fdtdx_plan_desc& d, int q, const std::function<void(fdtdx_detector_desc&)>& f) {
    std::memcpy(dt, d.detectors, sizeof(dt));
    f(dt[q]);
    d.detectors = dt;
  };
  malformed(sc, "a zero grid axis", [](fdtdx_plan_desc& d) { d.ny = 0; });
  malformed(sc, "negative steps", [](fdtdx_plan_desc& d) { d.num_steps = -1; });
  malformed(sc, "eps_comps 2", [](fdtdx_plan_desc& d) { d.eps_comps = 2; });
  malformed(sc, "mu_comps 4", [](fdtdx_plan_desc& d) { d.mu_comps = 4; });
  malformed(sc, "too many CPML slabs", [](fdtdx_plan_desc& d) { d.num_cpml = FDTDX_MAX_CPML + 1; });
  malformed(sc, "a negative source count", [](fdtdx_plan_desc& d) { d.num_sources = -1; });
  malformed(sc, "too many detectors", [](fdtdx_plan_desc& d) { d.num_detectors = FDTDX_MAX_DETECTORS + 1; });
  malformed(sc, "too many ADE boxes", [](fdtdx_plan_desc& d) { d.num_ade = FDTDX_MAX_ADE + 1; });
  malformed(sc, "too many walls", [](fdtdx_plan_desc& d) { d.num_walls = FDTDX_MAX_WALLS + 1; });
  malformed(sc, "a NULL CPML array", [](fdtdx_plan_desc& d) { d.cpml = nullptr; });
  malformed(sc, "a NULL detector array", [](fdtdx_plan_desc& d) { d.detectors = nullptr; });
  malformed(sc, "a NULL ADE array", [](fdtdx_plan_desc& d) { d.num_ade = 1; d.ade = nullptr; });
  malformed(sc, "a CPML range past the grid", [&](fdtdx_plan_desc& d) { cpml_edit(d, [](fdtdx_cpml_desc* c) { c[1].hi = kN + 3; }); });
  malformed(sc, "a CPML side 2", [&](fdtdx_plan_desc& d) { cpml_edit(d, [](fdtdx_cpml_desc* c) { c[4].side = 2; }); });
  malformed(sc, "a CPML axis 3", [&](fdtdx_plan_desc& d) { cpml_edit(d, [](fdtdx_cpml_desc* c) { c[2].axis = 3; }); });
  malformed(sc, "a missing CPML coefficient vector", [&](fdtdx_plan_desc& d) {
    cpml_edit(d, [](fdtdx_cpml_desc* c) { c[5].a_H = nullptr; });
  });
  malformed(sc, "a missing kappa vector", [&](fdtdx_plan_desc& d) {
    cpml_edit(d, [](fdtdx_cpml_desc* c) { c[3].kappa_is_one = 0; c[3].ikm1_E = nullptr; });
  });
  malformed(sc, "a source cell outside the grid", [&](fdtdx_plan_desc& d) { src_edit(d, [](fdtdx_source_desc& s) { s.lo[1] = kN; s.hi[1] = kN + 1; }); });
  malformed(sc, "a source component 3", [&](fdtdx_plan_desc& d) { src_edit(d, [](fdtdx_source_desc& s) { s.component = 3; }); });
  malformed(sc, "a source field 2", [&](fdtdx_plan_desc& d) { src_edit(d, [](fdtdx_source_desc& s) { s.field = 2; }); });
  malformed(sc, "a source kind 7", [&](fdtdx_plan_desc& d) { src_edit(d, [](fdtdx_source_desc& s) { s.kind = 7; }); });
  malformed(sc, "a source without a table", [&](fdtdx_plan_desc& d) { src_edit(d, [](fdtdx_source_desc& s) { s.table = nullptr; }); });
  static std::vector<float> amp3(3 * kT, 0.1f);  // a plane source's amplitude table is 3 * T floats
  malformed(sc, "a plane source without a profile", [&](fdtdx_plan_desc& d) {
    src_edit(d, [](fdtdx_source_desc& s) {
      s.table = amp3.data();
      s.kind = FDTDX_SRC_PLANE;
      s.lo[0] = 6; s.hi[0] = 7; s.lo[1] = 5; s.hi[1] = 11; s.lo[2] = 5; s.hi[2] = 11;
      s.target[0] = 1; s.target[1] = 2; s.src_comp[0] = 1; s.src_comp[1] = 2;
      s.profile = nullptr;
    });
  });
  malformed(sc, "a detector box outside the grid", [&](fdtdx_plan_desc& d) { det_edit(d, 2, [](fdtdx_detector_desc& q) { q.hi[2] = kN + 1; }); });
  malformed(sc, "an empty detector box", [&](fdtdx_plan_desc& d) { det_edit(d, 0, [](fdtdx_detector_desc& q) { q.hi[1] = q.lo[1]; }); });
  malformed(sc, "a detector kind 9", [&](fdtdx_plan_desc& d) { det_edit(d, 1, [](fdtdx_detector_desc& q) { q.kind = 9; }); });
  malformed(sc, "seven detector components", [&](fdtdx_plan_desc& d) { det_edit(d, 1, [](fdtdx_detector_desc& q) { q.num_components = 7; }); });
  malformed(sc, "a detector component id 6", [&](fdtdx_plan_desc& d) { det_edit(d, 1, [](fdtdx_detector_desc& q) { q.components[1] = 6; }); });
  malformed(sc, "a detector without time_to_idx", [&](fdtdx_plan_desc& d) { det_edit(d, 0, [](fdtdx_detector_desc& q) { q.time_to_idx = nullptr; }); });
  malformed(sc, "a phasor without frequencies", [&](fdtdx_plan_desc& d) { det_edit(d, 1, [](fdtdx_detector_desc& q) { q.num_freqs = 0; }); });
  malformed(sc, "a phasor without a window", [&](fdtdx_plan_desc& d) { det_edit(d, 1, [](fdtdx_detector_desc& q) { q.window = nullptr; }); });
  malformed(sc, "reduce_volume 3", [&](fdtdx_plan_desc& d) { det_edit(d, 2, [](fdtdx_detector_desc& q) { q.reduce_volume = 3; }); });
  malformed(sc, "an accumulating first detector", [&](fdtdx_plan_desc& d) { det_edit(d, 0, [](fdtdx_detector_desc& q) { q.accumulate = 1; }); });
  static fdtdx_wall_desc w[1];
  malformed(sc, "a wall kind 4", [&](fdtdx_plan_desc& d) { w[0] = {4, 0, 0}; d.num_walls = 1; d.walls = w; });
  malformed(sc, "a wall layer outside the grid", [&](fdtdx_plan_desc& d) { w[0] = {FDTDX_WALL_PEC, 1, kN}; d.num_walls = 1; d.walls = w; });
  malformed(sc, "periodic 2", [](fdtdx_plan_desc& d) { d.periodic[1] = 2; });
  malformed(sc, "a CPML slab on a periodic axis", [](fdtdx_plan_desc& d) { d.periodic[0] = 1; });
  malformed(sc, "complex_fields 2", [](fdtdx_plan_desc& d) { d.complex_fields = 2; });
  malformed(sc, "a Bloch phase on an axis that is not periodic", [](fdtdx_plan_desc& d) {
    d.complex_fields = 1;
    d.bloch_phase[2][0] = 0.5;
  });
  malformed(sc, "a non-finite Bloch phase", [](fdtdx_plan_desc& d) {
    d.num_cpml = 4;  // the x and y slabs only
    d.periodic[2] = 1;
    d.complex_fields = 1;
    d.bloch_phase[2][1] = NAN;
  });
  malformed(sc, "mirror 5", [](fdtdx_plan_desc& d) { d.mirror[0] = 5; });
  static std::vector<float> sig(64, 1.f);
  malformed(sc, "conductivity comps 2", [&](fdtdx_plan_desc& d) {
    d.sigma_E.comps = 2;
    for (int a = 0; a < 3; ++a) { d.sigma_E.lo[a] = 6; d.sigma_E.hi[a] = 10; }
    d.sigma_E.sigma = sig.data();
  });
  malformed(sc, "a conductivity box outside the grid", [&](fdtdx_plan_desc& d) {
    d.sigma_H.comps = 1;
    for (int a = 0; a < 3; ++a) { d.sigma_H.lo[a] = 6; d.sigma_H.hi[a] = 10; }
    d.sigma_H.hi[0] = kN + 2;
    d.sigma_H.sigma = sig.data();
  });
  malformed(sc, "conductivity without sigma", [&](fdtdx_plan_desc& d) {
    d.sigma_E.comps = 1;
    for (int a = 0; a < 3; ++a) { d.sigma_E.lo[a] = 6; d.sigma_E.hi[a] = 10; }
    d.sigma_E.sigma = nullptr;
  });
  malformed(sc, "a conductivity tensor without the full permittivity tensor", [&](fdtdx_plan_desc& d) {
    d.sigma_E.comps = 9;
    for (int a = 0; a < 3; ++a) { d.sigma_E.lo[a] = 6; d.sigma_E.hi[a] = 8; }
    d.sigma_E.sigma = sig.data();
  });
  static fdtdx_ade_desc ad[1];
  static std::vector<float> c123(3 * 64, 0.1f);
  auto ade_base = [&](fdtdx_plan_desc& d) {
    std::memset(ad, 0, sizeof(ad));
    for (int a = 0; a < 3; ++a) { ad[0].lo[a] = 6; ad[0].hi[a] = 10; }
    ad[0].num_poles = 1;
    ad[0].coef_comps = 1;
    ad[0].c1 = ad[0].c2 = ad[0].c3 = c123.data();
    d.num_ade = 1;
    d.ade = ad;
  };
  malformed(sc, "an ADE box outside the grid", [&](fdtdx_plan_desc& d) { ade_base(d); ad[0].lo[1] = -1; });
  malformed(sc, "an ADE box with no poles", [&](fdtdx_plan_desc& d) { ade_base(d); ad[0].num_poles = 0; });
  malformed(sc, "ADE coef_comps 2", [&](fdtdx_plan_desc& d) { ade_base(d); ad[0].coef_comps = 2; });
  malformed(sc, "ADE coupling 9 without the full tensor", [&](fdtdx_plan_desc& d) { ade_base(d); ad[0].coupling_comps = 9; });
  malformed(sc, "an ADE box without c2", [&](fdtdx_plan_desc& d) { ade_base(d); ad[0].c2 = nullptr; });
  malformed(sc, "record format 9", [](fdtdx_plan_desc& d) { d.record.format_E[2] = 9; });
  malformed(sc, "a record time filter starting past the end", [](fdtdx_plan_desc& d) {
    d.record.every_k = 4;
    d.record.start = kT + 5;
  });
  expect(fdtdx_plan_create(nullptr, nullptr) == FDTDX_EINVAL, "null descriptor");
  fdtdx_plan_destroy(nullptr);
  // the valid scene once more, after every rejected build
  valid_run(sc);
  std::printf("plan_asan: %d cases, %d failed\n", g_cases, g_failed);
  if (g_failed) return 1;
  std::printf("plan_asan: %d cases passed\n", g_cases);
  return 0;
}
