"""Interface recorder and its compression modules (reference: src/fdtdx/interfaces/recorder.py:14-199,
modules.py:98-161, time_filter.py:140-257).

The engine stores the PML interface layers of the reversible forward on the device, one slot per
kept step. ``DtypeConversion`` stores them in a narrow float type (round to nearest, ties to
even); ``LinearReconstructEveryK`` keeps every k-th step (and the last) and the reverse pass
reconstructs the others by linear interpolation. Modules apply in list order on compression and
in reverse order on decompression, so a time filter listed before a cast interpolates the
decoded float32 values, and one listed after it interpolates in the narrow type.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Sequence

# engine record formats (FDTDX_REC_* in include/fdtdx_hip.h)
FORMATS = {"float32": 0, "bfloat16": 1, "float16": 2, "float8_e4m3fn": 3, "float8_e4m3fnuz": 4, "float8_e5m2": 5,
           "float8_e5m2fnuz": 6}


def format_name(dtype) -> str:
    """The record format of a dtype given as a torch or NumPy dtype, a jax-style name or a string."""
    name = getattr(dtype, "name", None) or str(dtype)
    name = name.replace("torch.", "").replace("jnp.", "").replace("<class '", "").replace("'>", "")
    name = {"half": "float16", "float": "float32", "bf16": "bfloat16"}.get(name, name)
    if name.startswith("complex"):
        raise ValueError(f"DtypeConversion target dtype {name} is complex; the interface record is real")
    if name not in FORMATS:
        raise NotImplementedError(f"record dtype {name!r} is not supported by this engine "
                                  f"(supported: {', '.join(FORMATS)})")
    return name


@dataclass
class DtypeConversion:
    """modules.py:98-161: store every record entry whose name ("{pml}_{E|H}") contains none of the
    ``exclude_filter`` substrings as ``dtype``."""

    dtype: object
    exclude_filter: Sequence[str] = ()

    def __post_init__(self):
        self.format = format_name(self.dtype)
        self.exclude_filter = tuple(self.exclude_filter or ())

    def format_of(self, key: str) -> str:
        return "float32" if any(e in key for e in self.exclude_filter) else self.format


@dataclass
class LinearReconstructEveryK:
    """time_filter.py:140-257: keep steps start_recording_after, +k, ... and the last step."""

    k: int
    start_recording_after: int = 0

    def __post_init__(self):
        if self.k < 1:
            raise ValueError(f"k must be >= 1, got {self.k}")
        if self.start_recording_after < 0:
            raise ValueError("start_recording_after must be >= 0")


@dataclass
class Recorder:
    """recorder.py:14-199: the compression pipeline of the interface record."""

    modules: list = field(default_factory=list)

    def __post_init__(self):
        self.modules = list(self.modules)
        casts = [m for m in self.modules if isinstance(m, DtypeConversion)]
        filters = [m for m in self.modules if isinstance(m, LinearReconstructEveryK)]
        if len(casts) + len(filters) != len(self.modules):
            bad = [type(m).__name__ for m in self.modules if not isinstance(m, (DtypeConversion, LinearReconstructEveryK))]
            raise NotImplementedError(f"record modules {bad} are not supported by this engine")
        if len(casts) > 1 or len(filters) > 1:
            raise NotImplementedError("at most one DtypeConversion and one LinearReconstructEveryK are supported")

    def lower(self, pml_names: Sequence[str]) -> dict:
        """The engine's record descriptor for the PML objects `pml_names` (in object order)."""
        cast = next((m for m in self.modules if isinstance(m, DtypeConversion)), None)
        filt = next((m for m in self.modules if isinstance(m, LinearReconstructEveryK)), None)
        fmt_e = [FORMATS[cast.format_of(f"{n}_E")] if cast else 0 for n in pml_names]
        fmt_h = [FORMATS[cast.format_of(f"{n}_H")] if cast else 0 for n in pml_names]
        filter_first = 1
        if cast is not None and filt is not None:
            filter_first = int(self.modules.index(filt) < self.modules.index(cast))
        return dict(every_k=filt.k if filt else 1, start=filt.start_recording_after if filt else 0,
                    filter_first=filter_first, format_E=fmt_e, format_H=fmt_h)
