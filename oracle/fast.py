"""ctypes driver of the C oracle (oracle_c.c) -- TEST INFRASTRUCTURE, see fdtd_oracle.py.

`run(scene, T)` is a drop-in for fdtd_oracle.run(scene, T) on the scenes the benchmark configs
build (isotropic or diagonal inv_eps, non-magnetic, CPML, point and plane sources, ADE boxes,
exact-interpolation energy / field / flux / phasor detectors). It takes the tables of a
scenes.build_scene() Scene as they are, so both restatements share one set-up and differ only in
the time loop, which tests/test_oracle_c.py compares. Anything outside that scope raises
NotImplementedError. The loop runs on `threads` OpenMP threads (default: all cores).
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import fdtd_oracle as O

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_c.so"

_vp = C.c_void_p
_i32 = C.c_int32
_f64 = C.c_double


class _Slab(C.Structure):
    _fields_ = [("axis", _i32), ("lo", _i32), ("hi", _i32), ("kappa_one", _i32),
                ("bE", _vp), ("aE", _vp), ("ikE", _vp), ("bH", _vp), ("aH", _vp), ("ikH", _vp)]


class _Source(C.Structure):
    _fields_ = [("kind", _i32), ("field", _i32), ("comp", _i32), ("cell", _i32 * 3), ("table", _vp),
                ("axis", _i32), ("index", _i32), ("nterms", _i32), ("tcomp", _i32 * 2), ("tsrc", _i32 * 2),
                ("tsign", _f64 * 2), ("sign", _f64), ("profile", _vp), ("amp", _vp), ("mult", _vp * 3),
                ("mult_scalar", _f64)]


class _Ade(C.Structure):
    _fields_ = [("lo", _i32 * 3), ("hi", _i32 * 3), ("P", _i32), ("C", _i32), ("c1", _vp), ("c2", _vp), ("c3", _vp)]


class _Det(C.Structure):
    _fields_ = [("kind", _i32), ("lo", _i32 * 3), ("hi", _i32 * 3), ("time_to_idx", _vp), ("weight", _f64),
                ("sign", _f64), ("axis", _i32), ("ncomp", _i32), ("comps", _i32 * 6), ("nfreq", _i32),
                ("cos_t", _vp), ("sin_t", _vp), ("window", _vp), ("scale", _f64), ("state", _vp)]


class _Scene(C.Structure):
    _fields_ = [("nx", _i32), ("ny", _i32), ("nz", _i32), ("T", _i32), ("eps_comps", _i32), ("courant", _f64),
                ("inv_eps", _vp), ("nslab", _i32), ("slabs", C.POINTER(_Slab)), ("nsrc", _i32),
                ("src", C.POINTER(_Source)), ("nade", _i32), ("ade", C.POINTER(_Ade)), ("ndet", _i32),
                ("det", C.POINTER(_Det)), ("E", _vp), ("H", _vp), ("nthreads", _i32)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built (make -C {HERE})")
        _lib = C.CDLL(str(LIB))
        for name in ("oc_run_f32", "oc_run_f64"):
            getattr(_lib, name).argtypes = [C.POINTER(_Scene)]
            getattr(_lib, name).restype = C.c_int
    return _lib


_KIND = {"energy": 0, "field": 1, "flux": 2, "phasor": 3}


def _check(scene):
    if any(scene.periodic) or any(scene.mirror) or scene.walls:
        raise NotImplementedError("C oracle: periodic, mirror or wall faces")
    if scene.sigma_E is not None or scene.sigma_H is not None:
        raise NotImplementedError("C oracle: conductivity")
    if not isinstance(scene.inv_mu, float) or scene.inv_mu != 1.0:
        raise NotImplementedError("C oracle: magnetic materials")
    if scene.inv_eps.shape[0] not in (1, 3):
        raise NotImplementedError("C oracle: full-tensor permittivity")
    for d in scene.detectors:
        k = d["kind"]
        if k not in _KIND or not d.get("exact_interpolation", True) or d.get("inverse", False):
            raise NotImplementedError(f"C oracle: detector {d.get('name', k)!r}")
        if k in ("energy", "flux") and not d["reduce_volume"]:
            raise NotImplementedError("C oracle: unreduced energy / flux")
        if k == "flux" and d.get("keep_all_components", False):
            raise NotImplementedError("C oracle: keep_all_components flux")
        if k in ("field", "phasor") and d.get("reduce_volume", False):
            raise NotImplementedError("C oracle: reduced field / phasor")


def run(scene, T, threads=None):
    """The forward loop of fdtd_oracle.run(scene, T) from zero fields: returns (State with E and H,
    detector states in init_detector_state's layouts)."""
    _check(scene)
    lib = load()
    f = scene.dtype
    keep = []

    def ptr(a, dt=f):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    shape = scene.shape
    E = np.zeros((3, *shape), f)
    H = np.zeros((3, *shape), f)
    slabs = (_Slab * max(1, len(scene.slabs)))()
    for i, s in enumerate(scene.slabs):
        bE, aE, ikE, bH, aH, ikH = s.coeffs
        slabs[i] = _Slab(s.axis, s.lo, s.hi, int(np.all(ikE == 1) and np.all(ikH == 1)),
                         ptr(bE), ptr(aE), ptr(ikE), ptr(bH), ptr(aH), ptr(ikH))
        if np.all(ikE == 1) != np.all(ikH == 1):
            raise NotImplementedError("C oracle: kappa == 1 on one field only")
    srcs = (_Source * max(1, len(scene.sources)))()
    for i, s in enumerate(scene.sources):
        x = _Source()
        x.field = 0 if s["field"] == "E" else 1
        if s["kind"] == "dipole":
            x.kind, x.comp = 0, s["component"]
            x.cell = (_i32 * 3)(*s["cell"])
            x.table = ptr(s["table"])
        else:
            x.kind = 1
            ax = [a for a, sl in enumerate(s["slc"]) if sl != slice(None)]
            if len(ax) != 1:
                raise NotImplementedError("C oracle: plane source box")
            x.axis = ax[0]
            x.index = s["slc"][ax[0]].start
            x.nterms = len(s["terms"])
            for t, (comp, sc, ts) in enumerate(s["terms"]):
                x.tcomp[t], x.tsrc[t], x.tsign[t] = comp, sc, float(ts)
            x.sign = float(s["sign"])
            x.profile = ptr(s["profile"])
            x.amp = ptr(s["amp"])
            for comp in range(3):
                m = s["mult"].get(comp)
                if m is None or np.ndim(m) == 0:
                    x.mult[comp] = None
                    x.mult_scalar = float(1.0 if m is None else m)
                else:
                    x.mult[comp] = ptr(m)
        srcs[i] = x
    ades = (_Ade * 1)()
    nade = 0
    if scene.ade is not None:
        c1, c2, c3 = scene.ade
        nz_mask = np.any((c1 != 0) | (c3 != 0), axis=(0, 1))
        if np.any(nz_mask):
            idx = np.nonzero(nz_mask)
            lo = [int(v.min()) for v in idx]
            hi = [int(v.max()) + 1 for v in idx]
            sl = (slice(None), slice(None), *[slice(a, b) for a, b in zip(lo, hi)])
            ades[0] = _Ade((_i32 * 3)(*lo), (_i32 * 3)(*hi), c1.shape[0], c1.shape[1],
                           ptr(c1[sl]), ptr(c2[sl]), ptr(c3[sl]))
            nade = 1
    states = [O.init_detector_state(d, scene) for d in scene.detectors]
    dets = (_Det * max(1, len(scene.detectors)))()
    for i, d in enumerate(scene.detectors):
        x = _Det()
        x.kind = _KIND[d["kind"]]
        x.lo = (_i32 * 3)(*[sl.start for sl in d["slc"]])
        x.hi = (_i32 * 3)(*[sl.stop for sl in d["slc"]])
        x.time_to_idx = ptr(d["time_to_idx"], np.int32)
        x.weight = float(d.get("weight", 0.0))
        x.sign = float(d.get("sign", 1))
        x.axis = int(d.get("axis", 0))
        comps = d.get("components", ())
        x.ncomp = len(comps)
        for t, cn in enumerate(comps):
            x.comps[t] = O.COMPONENTS.index(cn)
        if d["kind"] == "phasor":
            x.nfreq = d["cos"].shape[0]
            x.cos_t, x.sin_t, x.window = ptr(d["cos"]), ptr(d["sin"]), ptr(d["window"])
            x.scale = float(d["scale"])
        x.state = states[i].ctypes.data
        dets[i] = x
    sc = _Scene(shape[0], shape[1], shape[2], int(T), scene.inv_eps.shape[0], float(scene.c), ptr(scene.inv_eps),
                len(scene.slabs), slabs, len(scene.sources), srcs, nade, ades, len(scene.detectors), dets,
                E.ctypes.data, H.ctypes.data, int(threads or os.cpu_count() or 1))
    rc = (lib.oc_run_f32 if f == np.float32 else lib.oc_run_f64)(C.byref(sc))
    if rc != 0:
        raise RuntimeError(f"C oracle returned {rc}")
    st = O.State(E, H, [], [])
    return st, states
