"""CPU oracle for the interface-record compression pipeline (TEST INFRASTRUCTURE).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module; the product never routes through it.

Restates, from the reference (paths relative to ``src/fdtdx``):
  * ``DtypeConversion`` (interfaces/modules.py:98-161): a cast of every record entry whose key
    does not contain an excluded substring to a narrow float type, and back on decompression.
    The casts are round-to-nearest-even; a value beyond the largest finite one becomes inf in
    the IEEE-like formats (bfloat16, float16, float8_e5m2) and NaN in the formats without inf
    (float8_e4m3fn, float8_e4m3fnuz, float8_e5m2fnuz); the "fnuz" formats have no negative zero.
  * ``LinearReconstructEveryK`` (interfaces/time_filter.py:140-257): keep the record of steps
    start, start+k, ... and the last step; reconstruct the others by linear interpolation
    between the two surrounding kept steps.
  * ``Recorder.compress`` / ``decompress`` (interfaces/recorder.py:70-199): modules applied in
    list order on compression and in reverse order on decompression.

Parity pinning (tests/test_record_codec.py): the codec against torch's own conversions to each
format (an independent implementation) over every representable value, every rounding midpoint
and random values; the time filter against the reference's unit-test expectations
(tests/unit/interfaces/test_time_filter.py, test_recorder.py).
"""

from __future__ import annotations

import numpy as np

# name -> (exponent bits, mantissa bits, bias, has_inf, fnuz)
FORMATS = {
    "float32": (8, 23, 127, True, False),
    "bfloat16": (8, 7, 127, True, False),
    "float16": (5, 10, 15, True, False),
    "float8_e4m3fn": (4, 3, 7, False, False),
    "float8_e4m3fnuz": (4, 3, 8, False, True),
    "float8_e5m2": (5, 2, 15, True, False),
    "float8_e5m2fnuz": (5, 2, 16, False, True),
}


def width(fmt: str) -> int:
    e, m, _, _, _ = FORMATS[fmt]
    return (1 + e + m) // 8


def _max_finite(fmt):
    e, m, b, has_inf, fnuz = FORMATS[fmt]
    if has_inf:
        top = (1 << e) - 2  # the all-ones exponent is inf / NaN
        frac = (1 << m) - 1
    elif fnuz:
        top = (1 << e) - 1  # only 0x80 is NaN
        frac = (1 << m) - 1
    else:  # e4m3fn: S.1111.111 is NaN
        top = (1 << e) - 1
        frac = (1 << m) - 2
    return top, frac


def decode(codes, fmt: str) -> np.ndarray:
    """Narrow float codes (unsigned integers) to float32."""
    e, m, b, has_inf, fnuz = FORMATS[fmt]
    if fmt == "float32":
        return np.asarray(codes, np.uint32).view(np.float32)
    c = np.asarray(codes, np.int64)
    nbits = 1 + e + m
    sign = (c >> (nbits - 1)) & 1
    ex = (c >> m) & ((1 << e) - 1)
    fr = c & ((1 << m) - 1)
    mag = np.where(ex == 0, fr.astype(np.float64) * 2.0 ** (1 - b - m),
                   (1.0 + fr.astype(np.float64) / (1 << m)) * np.exp2(ex.astype(np.float64) - b))
    out = np.where(sign == 1, -mag, mag)
    if has_inf:
        special = ex == (1 << e) - 1
        out = np.where(special & (fr == 0), np.where(sign == 1, -np.inf, np.inf), out)
        out = np.where(special & (fr != 0), np.nan, out)
    elif fnuz:
        out = np.where(c == (1 << (nbits - 1)), np.nan, out)
    else:
        out = np.where((ex == (1 << e) - 1) & (fr == (1 << m) - 1), np.nan, out)
    return out.astype(np.float32)


def encode(x, fmt: str) -> np.ndarray:
    """float32 to narrow float codes: round to nearest, ties to even, on the format's lattice
    extended past its largest finite value, then overflow to inf (IEEE-like) or NaN."""
    x = np.asarray(x, np.float32)
    bits = x.view(np.uint32).astype(np.int64)
    if fmt == "float32":
        return bits.astype(np.uint32)
    e, m, b, has_inf, fnuz = FORMATS[fmt]
    nbits = 1 + e + m
    sign = bits >> 31
    a = bits & 0x7FFFFFFF
    is_nan = a > 0x7F800000
    is_inf = a == 0x7F800000
    e32 = a >> 23
    f32 = a & 0x7FFFFF
    # value = mant * 2^(exp2 - 23) with mant the 24-bit significand (f32 subnormals: e32 = 0)
    mant = np.where(e32 == 0, f32, f32 | (1 << 23))
    exp2 = np.where(e32 == 0, -126, e32 - 127)
    emin = 1 - b
    # drop `sh` low bits so that the kept integer counts quanta of 2^(max(exp2, emin) - m)
    sh = 23 - m + np.maximum(emin - exp2, 0)
    sh = np.minimum(sh, 40)
    q = mant >> sh
    rem = mant - (q << sh)
    half = np.where(sh > 0, np.left_shift(np.int64(1), np.maximum(sh - 1, 0)), 0)
    up = (sh > 0) & ((rem > half) | ((rem == half) & ((q & 1) == 1)))
    q = q + up
    # q quanta of 2^(qexp - m), qexp = max(exp2, emin); renormalise a carry into the next binade
    qexp = np.maximum(exp2, emin)
    carry = q >= (1 << (m + 1))
    q = np.where(carry, q >> 1, q)
    qexp = qexp + carry
    normal = q >= (1 << m)
    ex_field = np.where(normal, qexp - emin + 1, 0)
    fr_field = np.where(normal, q - (1 << m), q)
    top, frac = _max_finite(fmt)
    over = (ex_field > top) | ((ex_field == top) & (fr_field > frac))
    mag_code = (ex_field << m) | fr_field
    signbit = sign << (nbits - 1)
    if has_inf:
        inf_code = ((1 << e) - 1) << m
        nan_code = inf_code | (1 << (m - 1))
        out = np.where(over | is_inf, signbit | inf_code, signbit | mag_code)
        out = np.where(is_nan, signbit | nan_code, out)
    elif fnuz:
        nan_code = 1 << (nbits - 1)
        out = np.where(mag_code == 0, 0, signbit | mag_code)  # no negative zero
        out = np.where(over | is_inf | is_nan, nan_code, out)
    else:
        nan_code = (1 << (nbits - 1)) - 1
        out = np.where(over | is_inf | is_nan, signbit | nan_code, signbit | mag_code)
    return out.astype(np.uint16 if nbits == 16 else np.uint8)


def cast(x, fmt: str) -> np.ndarray:
    """float32 -> fmt -> float32."""
    return decode(encode(x, fmt), fmt)


# ---------------------------------------------------------------- time filter (time_filter.py)
def time_filter_tables(T: int, k: int, start: int = 0):
    """LinearReconstructEveryK.init_shapes (time_filter.py:163-200): the kept steps and the
    step -> array index map built by the reference's forward fill (k - 1 rolls, the first k
    entries reset to 0 after each)."""
    save = list(range(start, T, k))
    if save[-1] != T - 1:
        save.append(T - 1)
    idx = np.zeros(T, np.int32)
    idx[save] = np.arange(len(save), dtype=np.int32)
    for _ in range(k - 1):
        rolled = np.roll(idx, 1)
        idx = np.where(idx == 0, rolled, idx)
        idx[:k] = 0
    return save, idx


def first_index(arr, val) -> int:
    """core/misc.py:202-215 (argmax of equality: 0 when absent)."""
    hits = np.nonzero(arr == val)[0]
    return int(hits[0]) if hits.size else 0


def interp_factor(t: int, arr_idx: int, time_to_arr) -> np.float32:
    """(t - t_prev) / (t_next - t_prev) as int32 / int32 true division in float32
    (time_filter.py:239-243)."""
    prev_t = first_index(time_to_arr, arr_idx)
    next_t = first_index(time_to_arr, arr_idx + 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.float32(np.float32(t - prev_t) / np.float32(next_t - prev_t))


class RecordPipeline:
    """Recorder (recorder.py:32-199) for modules given as ("every_k", k, start) and
    ("dtype", fmt, exclude) tuples, over record entries named "{face}_{E|H}"."""

    def __init__(self, modules, T: int):
        self.modules = list(modules)
        self.T = T
        filters = [m for m in self.modules if m[0] == "every_k"]
        if len(filters) > 1:
            raise NotImplementedError("one time filter at most")
        self.filter = filters[0] if filters else None
        if self.filter:
            self.save, self.tta = time_filter_tables(T, self.filter[1], self.filter[2])
        else:
            self.save, self.tta = list(range(T)), np.arange(T, dtype=np.int32)
        self.saved = set(self.save)
        self.data = {}  # array index -> {key: stored array}

    def _dtype_of(self, mod, key):
        return None if any(e in key for e in mod[2]) else mod[1]

    def compress(self, values: dict, n: int):
        if n not in self.saved:
            return
        out = {}
        for key, v in values.items():
            v = np.asarray(v, np.float32)
            for mod in self.modules:  # the time filter passes values through unchanged
                if mod[0] == "dtype":
                    fmt = self._dtype_of(mod, key)
                    if fmt is not None:
                        v = cast(v, fmt)
            out[key] = v
        self.data[int(self.tta[n])] = out

    def decompress(self, n: int) -> dict:
        """Modules in reverse order; the filter interpolates in the precision of the stage it
        sits at: after the casts back to float32 when it precedes the casts in the module list,
        in the narrow format (each operation rounded to it) when it follows them."""
        a = int(self.tta[n])
        if n in self.saved:
            return {k: v.copy() for k, v in self.data[a].items()}
        f = interp_factor(n, a, self.tta)
        prev, nxt = self.data[a], self.data[a + 1]
        pos = self.modules.index(self.filter)
        out = {}
        for key in prev:
            narrow = [self._dtype_of(m, key) for m in self.modules[:pos] if m[0] == "dtype"]
            narrow = [x for x in narrow if x is not None]
            if narrow:  # interpolation in the narrow format (the last cast before the filter)
                fmt = narrow[-1]
                ff = cast(np.float32(f), fmt)
                d = cast(nxt[key] - prev[key], fmt)
                out[key] = cast(prev[key] + cast(ff * d, fmt), fmt)
            else:
                out[key] = prev[key] + f * (nxt[key] - prev[key])
        return out
