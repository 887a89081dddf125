"""Simulation and gradient configuration (reference: src/fdtdx/config.py:15-307).

Only the uniform grid is supported by the engine; non-uniform rectilinear grids raise
NotImplementedError at placement.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Literal

import numpy as np

from . import constants


@dataclass(frozen=True)
class UniformGrid:
    """Uniform cubic grid policy (core/grid.py:13-60): ``spacing`` in metres."""

    spacing: float
    center: tuple = (0.0, 0.0, 0.0)

    def __post_init__(self):
        if self.spacing <= 0:
            raise ValueError(f"Uniform grid spacing must be positive, got {self.spacing}.")


@dataclass
class GradientConfig:
    """config.py:15-53. ``method`` is "reversible" (time-reversal adjoint, needs a Recorder) or
    "checkpointed" (forward checkpoints + recomputation, needs ``num_checkpoints``)."""

    method: Literal["reversible", "checkpointed"] = "reversible"
    recorder: object | None = None
    num_checkpoints: int | None = None
    num_checkpoints_reversible: int = 0
    # Device bytes the reversible path may spend on the interface record (not in the reference):
    # beyond it the record is regenerated chunk by chunk from full-state checkpoints, with
    # bit-identical gradients. None: half of the free device memory at the time of the forward.
    max_record_bytes: int | None = None

    def __post_init__(self):
        if self.method == "reversible" and self.recorder is None:
            raise Exception("Need Recorder in gradient config to compute reversible gradients")
        if self.method == "checkpointed" and self.num_checkpoints is None:
            raise Exception("Need Checkpoint Number in gradient config to compute checkpointed gradients")
        if self.num_checkpoints_reversible < 0:
            raise Exception("num_checkpoints_reversible must be >= 0")
        if self.max_record_bytes is not None and self.max_record_bytes < 0:
            raise Exception("max_record_bytes must be >= 0")


@dataclass
class SimulationConfig:
    """config.py:56-307 (uniform-grid subset)."""

    time: float
    grid: UniformGrid
    backend: str = "gpu"
    dtype: object = np.float32
    use_complex_fields: bool | None = None
    courant_factor: float = 0.99
    symmetry: tuple = (0, 0, 0)
    gradient_config: GradientConfig | None = None

    def __post_init__(self):
        if len(self.symmetry) != 3 or any(s not in (-1, 0, 1) for s in self.symmetry):
            raise ValueError(f"config.symmetry must be a length-3 tuple in {{-1, 0, +1}}, got {self.symmetry!r}")
        if not isinstance(self.grid, UniformGrid):
            raise NotImplementedError("only UniformGrid is supported by this engine")
        if np.dtype(self.dtype) != np.float32:
            raise NotImplementedError("the HIP engine computes in float32 only")

    @property
    def has_symmetry(self) -> bool:
        """config.py:92-112: some axis is reduced by a mirror plane (-1 electric, +1 magnetic)."""
        return any(s != 0 for s in self.symmetry)

    @property
    def courant_number(self) -> float:
        return self.courant_factor / math.sqrt(3)

    def uniform_spacing(self) -> float:
        return self.grid.spacing

    @property
    def resolved_grid(self):
        return None

    @property
    def has_nonuniform_grid(self) -> bool:
        return False

    @property
    def time_step_duration(self) -> float:
        # core/grid.py:495-511, uniform branch
        return (self.courant_factor / float(np.sqrt(3.0))) * self.grid.spacing / constants.c

    @property
    def time_steps_total(self) -> int:
        return round(self.time / self.time_step_duration)

    @property
    def max_travel_distance(self) -> float:
        return constants.c * self.time

    @property
    def only_forward(self) -> bool:
        return self.gradient_config is None

    @property
    def invertible_optimization(self) -> bool:
        if self.gradient_config is None:
            return False
        return self.gradient_config.recorder is not None
