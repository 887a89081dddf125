"""Binding of libfdtdx_hip.so (include/fdtdx_hip.h) and the lowering of placed objects to a plan.

The library is loaded with ctypes; every entry point takes plain pointers and sizes. Device memory
comes from torch tensors on the current ROCm device and the current torch stream is passed
through, so the engine composes with torch work on the same stream. There is no CPU fallback:
calling into the engine without the built library or without a GPU raises.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_float, c_int32, c_size_t, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfdtdx_hip.so"

FDTDX_MAX_CPML = 6
FDTDX_MAX_SOURCES = 8
FDTDX_MAX_ADE = 4
FDTDX_MAX_DETECTORS = 16
FDTDX_MAX_WALLS = 6

c_float_p = POINTER(c_float)
c_int32_p = POINTER(c_int32)


class CpmlDesc(ctypes.Structure):
    _fields_ = [("axis", c_int32), ("side", c_int32), ("lo", c_int32), ("hi", c_int32), ("kappa_is_one", c_int32),
                ("b_E", c_float_p), ("a_E", c_float_p), ("ikm1_E", c_float_p),
                ("b_H", c_float_p), ("a_H", c_float_p), ("ikm1_H", c_float_p)]


class SourceDesc(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("field", c_int32), ("component", c_int32), ("lo", c_int32 * 3),
                ("hi", c_int32 * 3), ("target", c_int32 * 2), ("src_comp", c_int32 * 2), ("sign", c_float * 2),
                ("table", c_float_p), ("profile", c_float_p)]


class AdeDesc(ctypes.Structure):
    _fields_ = [("lo", c_int32 * 3), ("hi", c_int32 * 3), ("num_poles", c_int32), ("coef_comps", c_int32),
                ("c1", c_float_p), ("c2", c_float_p), ("c3", c_float_p), ("coupling_comps", c_int32)]


class DetectorDesc(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("lo", c_int32 * 3), ("hi", c_int32 * 3), ("reduce_volume", c_int32),
                ("exact_interpolation", c_int32), ("num_components", c_int32), ("components", c_int32 * 6),
                ("axis", c_int32), ("sign", c_float), ("weight", c_float), ("num_freqs", c_int32),
                ("scale", c_float), ("time_to_idx", c_int32_p), ("cos_table", c_float_p),
                ("sin_table", c_float_p), ("window", c_float_p), ("accumulate", c_int32), ("inverse", c_int32)]


class WallDesc(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("axis", c_int32), ("index", c_int32)]


class RecordDesc(ctypes.Structure):
    _fields_ = [("every_k", c_int32), ("start", c_int32), ("filter_first", c_int32),
                ("format_E", c_int32 * FDTDX_MAX_CPML), ("format_H", c_int32 * FDTDX_MAX_CPML)]


class LossDesc(ctypes.Structure):
    _fields_ = [("comps", c_int32), ("lo", c_int32 * 3), ("hi", c_int32 * 3), ("sigma", POINTER(c_float))]


class PlanDesc(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("nz", c_int32), ("num_steps", c_int32), ("courant", c_float),
                ("eps_comps", c_int32), ("mu_comps", c_int32),
                ("num_cpml", c_int32), ("cpml", POINTER(CpmlDesc)),
                ("num_sources", c_int32), ("sources", POINTER(SourceDesc)),
                ("num_ade", c_int32), ("ade", POINTER(AdeDesc)),
                ("num_detectors", c_int32), ("detectors", POINTER(DetectorDesc)),
                ("num_walls", c_int32), ("walls", POINTER(WallDesc)), ("periodic", c_int32 * 3),
                ("sigma_E", LossDesc), ("sigma_H", LossDesc), ("record", RecordDesc), ("mirror", c_int32 * 3),
                ("complex_fields", c_int32), ("bloch_phase", (ctypes.c_double * 2) * 3)]


class Buffers(ctypes.Structure):
    _fields_ = [("E", c_void_p), ("H", c_void_p), ("inv_eps", c_void_p), ("inv_mu", c_void_p), ("psi", c_void_p),
                ("ade_state", c_void_p), ("detector_state", POINTER(c_void_p)), ("record", c_void_p),
                ("workspace", c_void_p)]


class AdjBuffers(ctypes.Structure):
    _fields_ = [("lam_E", c_void_p), ("lam_H", c_void_p), ("lam_psi", c_void_p), ("grad_inv_eps", c_void_p),
                ("detector_cotangent", POINTER(c_void_p))]


EXPORTED = ("fdtdx_plan_create", "fdtdx_plan_destroy", "fdtdx_psi_floats", "fdtdx_ade_floats",
            "fdtdx_record_floats_per_step", "fdtdx_record_slot_bytes", "fdtdx_record_slots", "fdtdx_record_slot_of",
            "fdtdx_workspace_bytes", "fdtdx_forward", "fdtdx_forward_timed",
            "fdtdx_reverse_adjoint", "fdtdx_reverse", "fdtdx_last_error", "fdtdx_batch_create", "fdtdx_batch_destroy",
            "fdtdx_batch_attach", "fdtdx_batch_detach", "fdtdx_batch_advance", "fdtdx_batch_step_index",
            "fdtdx_batch_flags", "fdtdx_stream_probe", "fdtdx_plan_schedule",
            "fdtdx_forward_noalloc", "fdtdx_stencil_probe")

FDTDX_MAX_BATCH = 64
SCHED_ONE, SCHED_QUAD, SCHED_FUSED, SCHED_MARCH = 0, 1, 2, 3
FLAG_NONFINITE = 1
FLAG_BLOWUP = 2

_lib = None


class EngineError(RuntimeError):
    pass


def load_library(path: Path | None = None):
    """Load libfdtdx_hip.so and declare the prototypes. Raises if the library is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise EngineError(f"{p} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(str(p))
    lib.fdtdx_plan_create.argtypes = [POINTER(PlanDesc), POINTER(c_void_p)]
    lib.fdtdx_plan_create.restype = c_int32
    lib.fdtdx_plan_destroy.argtypes = [c_void_p]
    lib.fdtdx_plan_destroy.restype = None
    for nm in ("fdtdx_psi_floats", "fdtdx_ade_floats", "fdtdx_record_floats_per_step", "fdtdx_record_slot_bytes",
               "fdtdx_record_slots", "fdtdx_workspace_bytes"):
        getattr(lib, nm).argtypes = [c_void_p, POINTER(c_size_t)]
        getattr(lib, nm).restype = c_int32
    lib.fdtdx_plan_schedule.argtypes = [c_void_p, POINTER(c_int32)]
    lib.fdtdx_plan_schedule.restype = c_int32
    lib.fdtdx_record_slot_of.argtypes = [c_void_p, c_int32, POINTER(c_int32)]
    lib.fdtdx_record_slot_of.restype = c_int32
    lib.fdtdx_forward.argtypes = [c_void_p, POINTER(Buffers), c_int32, c_int32, c_int32, c_void_p]
    lib.fdtdx_forward.restype = c_int32
    lib.fdtdx_forward_noalloc.argtypes = [c_void_p, POINTER(Buffers), c_int32, c_int32, c_int32, c_void_p]
    lib.fdtdx_forward_noalloc.restype = c_int32
    lib.fdtdx_forward_timed.argtypes = [c_void_p, POINTER(Buffers), c_int32, c_int32, c_int32, c_void_p,
                                        POINTER(ctypes.c_double)]  # double[2]
    lib.fdtdx_forward_timed.restype = c_int32
    lib.fdtdx_reverse_adjoint.argtypes = [c_void_p, POINTER(Buffers), POINTER(AdjBuffers), c_int32, c_int32,
                                          c_void_p]
    lib.fdtdx_reverse_adjoint.restype = c_int32
    lib.fdtdx_reverse.argtypes = [c_void_p, POINTER(Buffers), c_int32, c_int32, c_int32, c_void_p]
    lib.fdtdx_reverse.restype = c_int32
    lib.fdtdx_batch_create.argtypes = [c_int32, c_int32, c_float, POINTER(c_void_p)]
    lib.fdtdx_batch_create.restype = c_int32
    lib.fdtdx_batch_destroy.argtypes = [c_void_p]
    lib.fdtdx_batch_destroy.restype = None
    lib.fdtdx_batch_attach.argtypes = [c_void_p, c_int32, c_void_p, POINTER(Buffers), c_int32, c_int32, c_int32]
    lib.fdtdx_batch_attach.restype = c_int32
    lib.fdtdx_batch_detach.argtypes = [c_void_p, c_int32]
    lib.fdtdx_batch_detach.restype = c_int32
    lib.fdtdx_batch_advance.argtypes = [c_void_p, c_int32, c_int32, c_void_p]
    lib.fdtdx_batch_advance.restype = c_int32
    lib.fdtdx_batch_step_index.argtypes = [c_void_p, c_int32, c_int32_p]
    lib.fdtdx_batch_step_index.restype = c_int32
    lib.fdtdx_batch_flags.argtypes = [c_void_p, c_int32_p]
    lib.fdtdx_batch_flags.restype = c_int32
    lib.fdtdx_stream_probe.argtypes = [c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, ctypes.c_longlong,
                                       c_void_p]
    lib.fdtdx_stream_probe.restype = c_int32
    lib.fdtdx_stencil_probe.argtypes = [c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                        c_void_p]
    lib.fdtdx_stencil_probe.restype = c_int32
    lib.fdtdx_last_error.argtypes = []
    lib.fdtdx_last_error.restype = ctypes.c_char_p
    if path is None:
        _lib = lib
    return lib


def _check(lib, rc, what):
    if rc != 0:
        msg = lib.fdtdx_last_error().decode(errors="replace")
        if rc == 3:
            raise NotImplementedError(f"{what}: {msg}")
        raise EngineError(f"{what} failed ({rc}): {msg}")


def _fptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(c_float_p)


class Plan:
    """An immutable engine plan plus the metadata the host needs to lay out buffers.

    ``spec`` is the lowered description: dict with keys shape, T, courant, eps_comps, mu_comps,
    cpml (list), sources (list), ade (list), detectors (list). Host arrays referenced by the
    descriptors are kept alive until the plan is created (the library copies them)."""

    def __init__(self, spec: dict):
        self.spec = spec
        lib = load_library()
        self._lib = lib
        nx, ny, nz = spec["shape"]
        keep = []

        def f32(a):
            a = np.ascontiguousarray(np.asarray(a, np.float32))
            keep.append(a)
            return a

        cp = (CpmlDesc * max(1, len(spec["cpml"])))()
        for i, s in enumerate(spec["cpml"]):
            cp[i].axis, cp[i].side, cp[i].lo, cp[i].hi = s["axis"], s["side"], s["lo"], s["hi"]
            cp[i].kappa_is_one = int(s["kappa_is_one"])
            for nm in ("b_E", "a_E", "b_H", "a_H"):
                setattr(cp[i], nm, _fptr(f32(s[nm])))
            cp[i].ikm1_E = _fptr(f32(s["ikm1_E"]))
            cp[i].ikm1_H = _fptr(f32(s["ikm1_H"]))
        sd = (SourceDesc * max(1, len(spec["sources"])))()
        for i, s in enumerate(spec["sources"]):
            sd[i].kind = 0 if s["kind"] == "point" else 1
            sd[i].field = s["field"]
            sd[i].component = s.get("component", 0)
            for a in range(3):
                sd[i].lo[a], sd[i].hi[a] = s["lo"][a], s["hi"][a]
            if s["kind"] == "plane":
                for t in range(2):
                    sd[i].target[t], sd[i].src_comp[t], sd[i].sign[t] = s["target"][t], s["src_comp"][t], s["sign"][t]
                sd[i].profile = _fptr(f32(s["profile"]))
            sd[i].table = _fptr(f32(s["table"]))
        ad = (AdeDesc * max(1, len(spec["ade"])))()
        for i, s in enumerate(spec["ade"]):
            for a in range(3):
                ad[i].lo[a], ad[i].hi[a] = s["lo"][a], s["hi"][a]
            ad[i].num_poles, ad[i].coef_comps = s["num_poles"], s["coef_comps"]
            ad[i].coupling_comps = int(s.get("coupling_comps", 0))
            ad[i].c1, ad[i].c2, ad[i].c3 = _fptr(f32(s["c1"])), _fptr(f32(s["c2"])), _fptr(f32(s["c3"]))
        dd = (DetectorDesc * max(1, len(spec["detectors"])))()
        for i, s in enumerate(spec["detectors"]):
            dd[i].kind = s["kind"]
            for a in range(3):
                dd[i].lo[a], dd[i].hi[a] = s["lo"][a], s["hi"][a]
            dd[i].reduce_volume, dd[i].exact_interpolation = s["reduce"], s["exact"]
            comps = tuple(s["components"])
            dd[i].num_components = len(comps)
            for t, c in enumerate(comps):
                dd[i].components[t] = c
            dd[i].axis, dd[i].sign, dd[i].weight, dd[i].scale = s["axis"], s["sign"], s["weight"], s["scale"]
            dd[i].accumulate = int(s.get("accumulate", 0))
            dd[i].inverse = int(s.get("inverse", 0))
            tti = np.ascontiguousarray(np.asarray(s["time_to_idx"], np.int32))
            keep.append(tti)
            dd[i].time_to_idx = tti.ctypes.data_as(c_int32_p)
            if s["kind"] == 3:
                dd[i].num_freqs = s["cos"].shape[0]
                dd[i].cos_table, dd[i].sin_table = _fptr(f32(s["cos"])), _fptr(f32(s["sin"]))
                dd[i].window = _fptr(f32(s["window"]))
        walls = spec.get("walls", [])
        wd = (WallDesc * max(1, len(walls)))()
        for i, w in enumerate(walls):
            wd[i].kind, wd[i].axis, wd[i].index = w["kind"], w["axis"], w["index"]
        desc = PlanDesc(nx, ny, nz, spec["T"], spec["courant"], spec["eps_comps"], spec["mu_comps"],
                        len(spec["cpml"]), cp, len(spec["sources"]), sd, len(spec["ade"]), ad,
                        len(spec["detectors"]), dd, len(walls), wd,
                        (c_int32 * 3)(*[int(bool(p)) for p in spec.get("periodic", (0, 0, 0))]),
                        self._loss(spec.get("sigma_E"), keep), self._loss(spec.get("sigma_H"), keep),
                        self._record(spec.get("record"), len(spec["cpml"])),
                        (c_int32 * 3)(*[int(bool(m)) for m in spec.get("mirror", (0, 0, 0))]))
        desc.complex_fields = int(bool(spec.get("complex_fields", 0)))
        for a, ph in enumerate(spec.get("bloch_phase", ((0.0, 0.0),) * 3)):
            desc.bloch_phase[a][0], desc.bloch_phase[a][1] = float(ph[0]), float(ph[1])
        h = c_void_p()
        _check(lib, lib.fdtdx_plan_create(ctypes.byref(desc), ctypes.byref(h)), "fdtdx_plan_create")
        self._h = h
        self.psi_floats = self._query("fdtdx_psi_floats")
        self.ade_floats = self._query("fdtdx_ade_floats")
        self.record_floats_per_step = self._query("fdtdx_record_floats_per_step")
        self.record_slot_bytes = self._query("fdtdx_record_slot_bytes")
        self.record_slots = self._query("fdtdx_record_slots")
        T = spec["T"]
        so = np.full(T, -1, np.int32)
        v = c_int32()
        for n in range(T):
            _check(lib, lib.fdtdx_record_slot_of(h, n, ctypes.byref(v)), "fdtdx_record_slot_of")
            so[n] = v.value
        # slot_of[n]: slot of step n or -1; arr[n]: the slot at or before n (the reference's
        # step -> array index map, time_filter.py:180-198)
        self.slot_of = so
        kept = np.nonzero(so >= 0)[0]
        self.slot_step = kept.astype(np.int64)  # step of each slot
        arr = np.zeros(T, np.int64)
        if kept.size:
            pos = np.searchsorted(kept, np.arange(T), side="right") - 1
            arr = np.where(pos >= 0, so[kept[np.maximum(pos, 0)]], 0)
        self.slot_arr = arr
        self.workspace_bytes = self._query("fdtdx_workspace_bytes")

    @staticmethod
    def _record(rec, nslab):
        """Interface-record compression as an fdtdx_record_desc (None: every step in float32)."""
        d = RecordDesc()
        if rec is None:
            return d
        d.every_k, d.start, d.filter_first = int(rec["every_k"]), int(rec["start"]), int(rec["filter_first"])
        for s in range(nslab):
            d.format_E[s], d.format_H[s] = int(rec["format_E"][s]), int(rec["format_H"][s])
        return d

    @staticmethod
    def _loss(box, keep):
        """A conductivity box {lo, hi, sigma (comps, box)} as an fdtdx_loss_desc (None: none)."""
        if box is None:
            return LossDesc(0, (c_int32 * 3)(0, 0, 0), (c_int32 * 3)(0, 0, 0), None)
        sig = np.ascontiguousarray(np.asarray(box["sigma"], np.float32))
        keep.append(sig)
        return LossDesc(int(sig.shape[0]), (c_int32 * 3)(*box["lo"]), (c_int32 * 3)(*box["hi"]), _fptr(sig))

    def slot_after(self, n: int) -> int:
        """The last slot the reverse pass reads for step n: its own, or the next kept one."""
        return int(self.slot_of[n]) if self.slot_of[n] >= 0 else int(self.slot_arr[n]) + 1

    @property
    def schedule(self) -> int:
        """The forward sweep schedule of the plan (SCHED_ONE, SCHED_QUAD, SCHED_FUSED, SCHED_MARCH)."""
        out = c_int32()
        _check(self._lib, self._lib.fdtdx_plan_schedule(self._h, ctypes.byref(out)), "fdtdx_plan_schedule")
        return int(out.value)

    @property
    def quad(self) -> bool:
        """True when the forward sweeps update four z cells per lane."""
        return self.schedule == SCHED_QUAD

    def _query(self, nm):
        out = c_size_t()
        _check(self._lib, getattr(self._lib, nm)(self._h, ctypes.byref(out)), nm)
        return int(out.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.fdtdx_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h


def _ptr(t):
    return None if t is None else c_void_p(t.data_ptr())


def make_buffers(E, H, inv_eps, inv_mu, psi, ade_state, det_states, record, workspace):
    arr = (c_void_p * max(1, len(det_states)))(*[t.data_ptr() for t in det_states]) if det_states else None
    b = Buffers(_ptr(E), _ptr(H), _ptr(inv_eps), _ptr(inv_mu), _ptr(psi), _ptr(ade_state),
                ctypes.cast(arr, POINTER(c_void_p)) if arr is not None else None, _ptr(record), _ptr(workspace))
    return b, arr


def current_stream_handle(device=None):
    import torch

    return c_void_p(torch.cuda.current_stream(device).cuda_stream)


def forward_noalloc(plan: Plan, buffers, t0: int, t1: int, record: bool, stream):
    """fdtdx_forward_noalloc: the forward with its workspace in the library's static device scratch
    (buffers.workspace is ignored; nothing is allocated)."""
    lib = plan._lib
    _check(lib, lib.fdtdx_forward_noalloc(plan.handle, ctypes.byref(buffers), int(t0), int(t1), int(bool(record)),
                                          stream), "fdtdx_forward_noalloc")


def forward(plan: Plan, buffers, t0: int, t1: int, record: bool, stream):
    lib = plan._lib
    _check(lib, lib.fdtdx_forward(plan.handle, ctypes.byref(buffers), int(t0), int(t1), int(bool(record)), stream),
           "fdtdx_forward")


def forward_timed(plan: Plan, buffers, t0: int, t1: int, record: bool, stream) -> tuple[float, float]:
    """fdtdx_forward plus HIP-event timing of the field-update launches; returns the summed ms of
    the E sweeps and of the H sweeps (a single-pass schedule reports (total, 0.0))."""
    lib = plan._lib
    ms = (ctypes.c_double * 2)()
    _check(lib, lib.fdtdx_forward_timed(plan.handle, ctypes.byref(buffers), int(t0), int(t1), int(bool(record)),
                                        stream, ms), "fdtdx_forward_timed")
    return float(ms[0]), float(ms[1])


def reverse(plan: Plan, buffers, t1: int, t0: int, reset_fields: bool, record_detectors: bool, stream):
    """fdtdx_reverse: the time-reversed run from step t1 back to t0 (full_backward)."""
    lib = plan._lib
    flags = (1 if reset_fields else 0) | (2 if record_detectors else 0)
    _check(lib, lib.fdtdx_reverse(plan.handle, ctypes.byref(buffers), int(t1), int(t0), flags, stream), "fdtdx_reverse")


def reverse_adjoint(plan: Plan, buffers, lam_E, lam_H, lam_psi, grad, cotangents, t1: int, t0: int, stream):
    lib = plan._lib
    arr = (c_void_p * max(1, len(cotangents)))(*[0 if c is None else c.data_ptr() for c in cotangents])
    adj = AdjBuffers(_ptr(lam_E), _ptr(lam_H), _ptr(lam_psi), _ptr(grad), ctypes.cast(arr, POINTER(c_void_p)))
    _check(lib, lib.fdtdx_reverse_adjoint(plan.handle, ctypes.byref(buffers), ctypes.byref(adj), int(t1), int(t0),
                                          stream), "fdtdx_reverse_adjoint")


class Batch:
    """A multi-instance batch (fdtdx_batch_*): up to FDTDX_MAX_BATCH solves of one material class
    advanced together, each kernel launch covering every attached instance. The plan and the buffer
    tensors of an attached solve are kept referenced here until it is detached."""

    def __init__(self, eps_comps: int = 1, mu_comps: int = 0, blowup_limit: float = 0.0):
        lib = load_library()
        self._lib = lib
        h = c_void_p()
        _check(lib, lib.fdtdx_batch_create(int(eps_comps), int(mu_comps), float(blowup_limit), ctypes.byref(h)),
               "fdtdx_batch_create")
        self._h = h
        self._held = {}

    def attach(self, slot: int, plan: Plan, buffers, keep, t0: int, t1: int, record: bool = False):
        _check(self._lib, self._lib.fdtdx_batch_attach(self._h, int(slot), plan.handle, ctypes.byref(buffers),
                                                       int(t0), int(t1), int(bool(record))), "fdtdx_batch_attach")
        self._held[int(slot)] = (plan, buffers, keep)

    def detach(self, slot: int):
        _check(self._lib, self._lib.fdtdx_batch_detach(self._h, int(slot)), "fdtdx_batch_detach")
        self._held.pop(int(slot), None)

    def advance(self, steps: int, stream, graph: bool = True):
        _check(self._lib, self._lib.fdtdx_batch_advance(self._h, int(steps), int(bool(graph)), stream),
               "fdtdx_batch_advance")

    def step_index(self, slot: int) -> int:
        n = c_int32()
        _check(self._lib, self._lib.fdtdx_batch_step_index(self._h, int(slot), ctypes.byref(n)),
               "fdtdx_batch_step_index")
        return int(n.value)

    def flags(self) -> np.ndarray:
        """Per-slot health flags (FLAG_*) as of the most recently completed chunk."""
        out = (c_int32 * FDTDX_MAX_BATCH)()
        _check(self._lib, self._lib.fdtdx_batch_flags(self._h, out), "fdtdx_batch_flags")
        return np.frombuffer(out, dtype=np.int32).copy()

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.fdtdx_batch_destroy(h)
            self._h = None
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
