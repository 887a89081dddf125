"""Small host-side value types: WaveCharacter (core/wavelength.py:5-65), OnOffSwitch
(core/switch.py:7-222) and the detector apodization windows (core/window.py:58-120)."""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import constants


@dataclass(frozen=True)
class WaveCharacter:
    period: float | None = None
    wavelength: float | None = None
    frequency: float | None = None
    phase_shift: float = 0.0

    def __post_init__(self):
        if sum([self.period is not None, self.frequency is not None, self.wavelength is not None]) != 1:
            raise Exception("Need to set exactly one of Period, Frequency or Wavelength in WaveCharacter")

    def get_period(self) -> float:
        if self.period is not None:
            return self.period
        if self.wavelength is not None:
            return self.wavelength / constants.c
        return 1.0 / self.frequency

    def get_wavelength(self) -> float:
        if self.wavelength is not None:
            return self.wavelength
        if self.period is not None:
            return self.period * constants.c
        return constants.c / self.frequency

    def get_frequency(self) -> float:
        if self.frequency is not None:
            return self.frequency
        if self.period is not None:
            return 1.0 / self.period
        return constants.c / self.wavelength


@dataclass(frozen=True)
class OnOffSwitch:
    start_time: float | None = None
    start_after_periods: float | None = None
    end_time: float | None = None
    end_after_periods: float | None = None
    on_for_time: float | None = None
    on_for_periods: float | None = None
    period: float | None = None
    fixed_on_time_steps: tuple | list | None = None
    is_always_off: bool = False
    interval: int = 1

    @property
    def is_default_always_on(self) -> bool:
        return (not self.is_always_off and self.fixed_on_time_steps is None and self.interval == 1
                and self.start_time is None and self.start_after_periods is None and self.end_time is None
                and self.end_after_periods is None and self.on_for_time is None and self.on_for_periods is None
                and self.period is None)

    def calculate_on_list(self, num_total_time_steps: int, time_step_duration: float) -> list[bool]:
        if self.fixed_on_time_steps is not None:
            on = [False] * num_total_time_steps
            for t in self.fixed_on_time_steps:
                on[t] = True
            return on
        return [self.is_on_at_time_step(t, time_step_duration) and t % self.interval == 0
                for t in range(num_total_time_steps)]

    def calculate_time_step_to_on_arr_idx(self, num_total_time_steps: int, time_step_duration: float) -> list[int]:
        on = self.calculate_on_list(num_total_time_steps, time_step_duration)
        out, c = [-1] * num_total_time_steps, 0
        for t in range(num_total_time_steps):
            if on[t]:
                out[t] = c
                c += 1
        return out

    def is_on_at_time_step(self, time_step: int, time_step_duration: float) -> bool:
        """core/switch.py:108-207."""
        if self.is_always_off:
            return False
        need_period = any(x is not None for x in (self.start_after_periods, self.end_after_periods, self.on_for_periods))
        if need_period and self.period is None:
            raise Exception("Need to specify period!")
        start_time, end_time, on_for_time = self.start_time, self.end_time, self.on_for_time
        n_start = sum([start_time is not None, self.start_after_periods is not None,
                       on_for_time is not None and end_time is not None,
                       self.on_for_periods is not None and end_time is not None,
                       on_for_time is not None and self.end_after_periods is not None,
                       self.on_for_periods is not None and self.end_after_periods is not None])
        if n_start > 1:
            raise Exception("Invalid start time specification!")
        if n_start == 0:
            start_time = 0
        n_end = sum([end_time is not None, self.end_after_periods is not None,
                     on_for_time is not None and start_time is not None,
                     self.on_for_periods is not None and start_time is not None,
                     on_for_time is not None and self.start_after_periods is not None,
                     self.on_for_periods is not None and self.start_after_periods is not None])
        if n_end > 1:
            raise Exception("Invalid end time specification!")
        if n_end == 0:
            end_time = math.inf
        if self.start_after_periods is not None:
            start_time = self.start_after_periods * self.period
        if self.end_after_periods is not None:
            end_time = self.end_after_periods * self.period
        if self.on_for_periods is not None:
            on_for_time = self.on_for_periods * self.period
        if start_time is None and on_for_time is not None:
            start_time = end_time - on_for_time
        if end_time is None and on_for_time is not None:
            end_time = start_time + on_for_time
        t = time_step * time_step_duration
        return bool(start_time <= t <= end_time)


# ---------------------------------------------------------------- apodization windows
# Evaluated in float32 with the operation order of the reference's weak-typed JAX expressions,
# on the float32 sample times n * dt.
class TemporalWindow:
    """window.py:58-72: a carrier-free window w(t) >= 0 applied to each phasor sample."""

    def get_window(self, time) -> np.ndarray:  # pragma: no cover - abstract
        raise NotImplementedError()


def gaussian_envelope(time, center, sigma):
    """window.py:16-22: exp(-(t - center)^2 / (2 sigma^2))."""
    f32 = np.float32
    d = (np.asarray(time, f32) - f32(center)).astype(f32)
    return np.exp((-(d * d)) / f32(2.0 * sigma**2)).astype(f32)


def tukey_envelope(time, start, end, alpha=0.5):
    """window.py:34-55: flat top 1 over [start, end] with cosine tapers over a fraction alpha,
    zero outside."""
    f32 = np.float32
    x = ((np.asarray(time, f32) - f32(start)) / f32(end - start)).astype(f32)
    in_range = (x >= 0) & (x <= 1)
    if alpha <= 0.0:
        return np.where(in_range, f32(1), f32(0)).astype(f32)
    half = alpha / 2.0
    left = (f32(0.5) * (f32(1) + np.cos(f32(np.pi) * (x / f32(half) - f32(1))))).astype(f32)
    right = (f32(0.5) * (f32(1) + np.cos(f32(np.pi) * ((x - f32(1)) / f32(half) + f32(1))))).astype(f32)
    w = np.where(x < half, left, np.where(x > 1.0 - half, right, f32(1))).astype(f32)
    return np.where(in_range, w, f32(0)).astype(f32)


@dataclass(frozen=True)
class GaussianWindow(TemporalWindow):
    """window.py:75-94."""

    center_time: float
    sigma_time: float

    def __post_init__(self):
        if not self.sigma_time > 0:
            raise ValueError(f"sigma_time must be positive, got {self.sigma_time}")

    def get_window(self, time):
        return gaussian_envelope(time, self.center_time, self.sigma_time)


@dataclass(frozen=True)
class TukeyWindow(TemporalWindow):
    """window.py:97-120 (alpha = 0: rectangular, alpha = 1: Hann)."""

    start_time: float
    end_time: float
    alpha: float = 0.5

    def __post_init__(self):
        if not self.end_time > self.start_time:
            raise ValueError(f"end_time must exceed start_time, got {self.start_time} and {self.end_time}")
        if not 0.0 <= self.alpha <= 1.0:
            raise ValueError(f"alpha must lie in [0, 1], got {self.alpha}")

    def get_window(self, time):
        return tukey_envelope(time, self.start_time, self.end_time, self.alpha)
