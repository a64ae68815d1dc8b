"""GPU permanent calculators: drop-in counterparts of theboss's calculators whose
arithmetic runs in the gfx950 kernels of libtheboss_amd.so.

Base-class semantics follow theboss's BSPermanentCalculatorBase
(theboss/boson_sampling_utilities/permanent_calculators/bs_permanent_calculator_base.py:22-72):
the matrix and states are stored by reference (None -> empty list), so in-place edits
such as `calc.matrix *= sqrt(eta)` are seen by the next computation; the shape check
raises AttributeError (:179-180). Each compute call gathers only the submatrix it needs
on the host and ships that to the device; no device state is kept in the objects, so
they deepcopy and pickle like the reference's. The all-pairs entry point keeps a small
module-level cache of device copies keyed by content (the matrix and the fixed side of
the pairs), which lock-step samplers pass unchanged on every round.
"""
import hashlib
from collections import OrderedDict
from typing import List, Optional, Sequence

import numpy as np
from numpy import complex128

from .. import _device as D
from .. import _lib as L
from .. import ops
from ..states import as_occupation, compact_batch, effective_matrix, layer_rows
from .bs_permanent_calculator_interface import BSPermanentCalculatorInterface


class BSPermanentCalculatorBase(BSPermanentCalculatorInterface):
    def __init__(self, matrix, input_state: Optional[Sequence[int]] = None,
                 output_state: Optional[Sequence[int]] = None) -> None:
        if output_state is None:
            output_state = list()
        if input_state is None:
            input_state = list()
        self._matrix = matrix
        self._input_state = input_state
        self._output_state = output_state

    @property
    def matrix(self):
        return self._matrix

    @matrix.setter
    def matrix(self, matrix) -> None:
        self._matrix = matrix

    @property
    def input_state(self):
        return self._input_state

    @input_state.setter
    def input_state(self, input_state) -> None:
        self._input_state = input_state

    @property
    def output_state(self):
        return self._output_state

    @output_state.setter
    def output_state(self, output_state) -> None:
        self._output_state = output_state

    def _can_calculation_be_performed(self) -> bool:
        """bs_permanent_calculator_base.py:61-72: square matrix, equal state lengths, no
        more modes than columns."""
        return (
            len(self._matrix) == len(self._matrix[0])
            and len(self._output_state) == len(self._input_state)
            and len(self._output_state) <= len(self._matrix[0])
        )

    def _checked_states(self):
        if self._matrix is None or not self._can_calculation_be_performed():
            raise AttributeError("matrix and states have incompatible sizes")
        return as_occupation(self._input_state), as_occupation(self._output_state)

    def compute_permanent_batch(self, input_states, output_states) -> np.ndarray:
        """Permanents for many (input, output) pairs on the same matrix in one launch
        (no reference counterpart; the reference loops over compute_permanent)."""
        raise NotImplementedError

    def compute_permanent_batch_arrays(self, S: np.ndarray, Tm: np.ndarray) -> np.ndarray:
        """compute_permanent_batch for pre-stacked B x m occupation arrays."""
        return self.compute_permanent_batch(list(S), list(Tm))


def _glynn_on_effective(matrices: List[np.ndarray], precision: int) -> np.ndarray:
    """Glynn permanents of equal-size square effective matrices (one launch)."""
    T = D.require_gpu()
    n = matrices[0].shape[0]
    A = D.to_device_c128(np.stack(matrices).reshape(len(matrices), n, n))
    out = ops.glynn_batch(A, precision)
    return out.cpu().numpy()


class GPUGlynnGrayPermanentCalculator(BSPermanentCalculatorBase):
    """Glynn's formula in Gray order on the GPU
    (glynn_gray_permanent_calculator.py:41-84). Empty effective matrix -> 1 (:52-53)."""

    precision = L.TB_PREC_FP64

    def compute_permanent(self) -> complex128:
        s, t = self._checked_states()
        if s.sum() == 0 or t.sum() == 0:
            return complex128(1)
        if s.sum() != t.sum():
            raise AttributeError("effective matrix is not square (sum(input) != sum(output))")
        M = effective_matrix(self._matrix, s, t)
        return complex128(_glynn_on_effective([M], self.precision)[0])

    def compute_permanent_batch(self, input_states, output_states) -> np.ndarray:
        out = np.ones(len(input_states), dtype=np.complex128)
        groups = {}
        for idx, (si, ti) in enumerate(zip(input_states, output_states)):
            s, t = as_occupation(si), as_occupation(ti)
            if s.sum() == 0 or t.sum() == 0:
                continue
            if s.sum() != t.sum():
                raise AttributeError("effective matrix is not square")
            groups.setdefault(int(s.sum()), []).append((idx, effective_matrix(self._matrix, s, t)))
        for n, items in groups.items():
            vals = _glynn_on_effective([M for _, M in items], self.precision)
            for (idx, _), v in zip(items, vals):
                out[idx] = v
        return out


class GPUGlynnGrayFP32PermanentCalculator(GPUGlynnGrayPermanentCalculator):
    """Glynn with FP32 column sums and products, FP64 accumulation of the alternating
    sum (no reference counterpart: a precision/throughput variant)."""

    precision = L.TB_PREC_FP32_COLSUM


_OUTER_CACHE: "OrderedDict" = OrderedDict()
_OUTER_CACHE_MAX = 32


def _content_key(tag: str, dev, a: np.ndarray):
    a = np.ascontiguousarray(a)
    return (tag, str(dev), a.dtype.str, a.shape, hashlib.blake2b(a.tobytes(), digest_size=16).digest())


def _compact_on_device(T, states: np.ndarray, dev):
    """compact_batch on the device: (modes, multiplicities) B x W int32 tensors, W the
    largest number of occupied modes of a row (counted on the host), each row's occupied
    modes in mode order, padded with mode 0 of multiplicity 0. One upload and a scatter,
    no synchronisation."""
    W = max(1, int((states > 0).sum(axis=1).max())) if len(states) else 1
    t = T.from_numpy(np.ascontiguousarray(states, dtype=np.int32)).to(dev)
    B, m = t.shape
    occ = t > 0
    slot = T.where(occ, T.cumsum(occ, dim=1, dtype=T.int64) - 1, T.full_like(t, W, dtype=T.int64))
    cols = T.arange(m, device=dev, dtype=T.int32).expand(B, m)
    modes = T.zeros((B, W + 1), dtype=T.int32, device=dev).scatter_(1, slot, cols)
    mult = T.zeros((B, W + 1), dtype=T.int32, device=dev).scatter_(1, slot, t)
    return [modes[:, :W].contiguous(), mult[:, :W].contiguous()]


def _cached(key, make):
    """Least-recently-used cache of device copies (content-keyed, so in-place edits of a
    matrix or state array give a new key)."""
    v = _OUTER_CACHE.get(key)
    if v is None:
        v = make()
        _OUTER_CACHE[key] = v
        while len(_OUTER_CACHE) > _OUTER_CACHE_MAX:
            _OUTER_CACHE.popitem(last=False)
    else:
        _OUTER_CACHE.move_to_end(key)
    return v


class _GuanGPUCalculator(BSPermanentCalculatorBase):
    """Multiplicity-aware Guan-code calculators (Ryser, Chin-Huh) on the GPU."""

    formula = L.TB_FORMULA_RYSER

    def _launch(self, S: np.ndarray, Tm: np.ndarray) -> np.ndarray:
        """One or two launches for stacked occupations S, Tm (B x m) on the shared matrix.
        Items whose output side has fewer Guan codes are run transposed
        (perm(U_ST) = perm(U^T_TS)), so the kernel always enumerates the smaller side."""
        U = np.asarray(self._matrix, dtype=np.complex128)
        B = S.shape[0]
        res = np.empty(B, dtype=np.complex128)
        codes_s = np.prod(S.astype(np.float64) + 1, axis=1)
        codes_t = np.prod(Tm.astype(np.float64) + 1, axis=1)
        flip = codes_t < codes_s
        dev = D.device()
        for transposed in (False, True):
            idx = np.nonzero(flip == transposed)[0]
            if idx.size == 0:
                continue
            enum_side, prod_side = (Tm[idx], S[idx]) if transposed else (S[idx], Tm[idx])
            in_modes, in_mult = compact_batch(enum_side)
            out_modes, out_mult = compact_batch(prod_side)
            if in_modes.shape[1] > L.TB_GUAN_MAX_COLS or out_modes.shape[1] > L.TB_GUAN_MAX_ROWS \
                    or np.max(np.where(transposed, codes_t[idx], codes_s[idx])) > 2.0 ** 28:
                return None
            M = U.T if transposed else U
            r = ops.guan_perm_batch(D.to_device_c128(M, dev), D.to_device_i32(out_modes, dev),
                                    D.to_device_i32(out_mult, dev), D.to_device_i32(in_modes, dev),
                                    D.to_device_i32(in_mult, dev), self.formula)
            res[idx] = r.cpu().numpy()
        return res

    def compute_permanent_outer(self, S: np.ndarray, Tm: np.ndarray):
        """All pairs at once: a (len(Tm), len(S)) complex128 device tensor holding
        perm(U[Tm[j], S[i]]) at [j, i], for S and Tm rows that all hold the same number
        k >= 1 of particles. The S side is compacted once (cached by content), the Tm side
        on the device, and one tb_guan_perm_outer launch walks every pair, each over the
        side with fewer Guan codes (as compute_permanent_batch_arrays chooses). None when
        a side needs the Glynn fallback."""
        S = np.asarray(S, dtype=np.int64)
        Tm = np.asarray(Tm, dtype=np.int64)
        k = int(S[0].sum()) if len(S) else 0
        if k < 1 or np.any(S.sum(axis=1) != k) or np.any(Tm.sum(axis=1) != k):
            raise AttributeError("outer batch needs rows with one common particle number >= 1")
        T = D.require_gpu()
        dev = D.device()
        codes_s, cs = _cached(_content_key("S", dev, S), lambda: (
            np.prod(S.astype(np.float64) + 1, axis=1),
            [D.to_device_i32(a, dev) for a in compact_batch(S)]))
        codes_t = np.prod(Tm.astype(np.float64) + 1, axis=1)
        ct = _compact_on_device(T, Tm, dev)
        if max(cs[0].shape[1], ct[0].shape[1]) > min(L.TB_GUAN_MAX_COLS, L.TB_GUAN_MAX_ROWS) \
                or max(codes_s.max(), codes_t.max()) > 2.0 ** 28:
            return None
        U = np.asarray(self._matrix, dtype=np.complex128)
        Ud = _cached(_content_key("U", dev, U), lambda: D.to_device_c128(U, dev))
        return ops.guan_perm_outer(Ud, ct[0], ct[1], cs[0], cs[1], self.formula)

    def compute_permanent_layer(self, S: np.ndarray, prefixes: np.ndarray):
        """A GCC-A layer's pairs: a (len(prefixes) * m, len(S)) complex128 device tensor
        holding perm(U[prefix_p + e_o, S[i]]) at [p * m + o, i], for S rows of k >= 1
        particles and prefixes of k - 1. The output rows are built in compact form from the
        compacted prefixes (each prefix's modes, then mode o: one more particle where the
        prefix holds o, else a new slot), uploaded once, and walked by one
        tb_guan_perm_outer launch. A row's slots are then not in mode order, which changes
        the walk's rounding only. None when a side needs the Glynn fallback."""
        S = np.asarray(S, dtype=np.int64)
        Pf = np.asarray(prefixes, dtype=np.int64)
        k = int(S[0].sum()) if len(S) else 0
        if k < 1 or np.any(S.sum(axis=1) != k) or Pf.ndim != 2 or Pf.shape[1] != S.shape[1] \
                or np.any(Pf.sum(axis=1) != k - 1):
            raise AttributeError("a layer needs S rows of k >= 1 particles and prefixes of k - 1")
        D.require_gpu()
        dev = D.device()
        codes_s, cs = _cached(_content_key("S", dev, S), lambda: (
            np.prod(S.astype(np.float64) + 1, axis=1),
            [D.to_device_i32(a, dev) for a in compact_batch(S)]))
        rows = layer_rows(Pf)  # modes, multiplicities: 2 x (P m) x (W + 1)
        codes_t = np.prod(rows[1].astype(np.float64) + 1, axis=1)
        if max(cs[0].shape[1], rows.shape[2]) > min(L.TB_GUAN_MAX_COLS, L.TB_GUAN_MAX_ROWS) \
                or max(codes_s.max(), codes_t.max()) > 2.0 ** 28:
            return None
        td = D.to_device_i32(rows, dev)
        U = np.asarray(self._matrix, dtype=np.complex128)
        Ud = _cached(_content_key("U", dev, U), lambda: D.to_device_c128(U, dev))
        return ops.guan_perm_outer(Ud, td[0], td[1], cs[0], cs[1], self.formula)

    def compute_permanent(self) -> complex128:
        s, t = self._checked_states()
        if len(s) == 0:
            raise AttributeError("empty occupation vectors")
        return complex128(self.compute_permanent_batch([s], [t])[0])

    def compute_permanent_batch(self, input_states, output_states) -> np.ndarray:
        pairs = [(as_occupation(s), as_occupation(t)) for s, t in zip(input_states, output_states)]
        if not pairs:
            return np.zeros(0, dtype=np.complex128)
        if any(len(p[0]) != len(pairs[0][0]) or len(p[1]) != len(p[0]) for p in pairs):
            raise AttributeError("occupation vectors of a batch must share one length")
        return self.compute_permanent_batch_arrays(np.stack([p[0] for p in pairs]),
                                                   np.stack([p[1] for p in pairs]))

    def compute_permanent_batch_arrays(self, S: np.ndarray, Tm: np.ndarray) -> np.ndarray:
        S = np.asarray(S, dtype=np.int64)
        Tm = np.asarray(Tm, dtype=np.int64)
        res = self._launch(S, Tm)
        if res is None:
            pairs = list(zip(S, Tm))
            # Too many distinct modes for the mixed-radix kernel: expand multiplicities
            # and use the Glynn kernel on the effective matrices (same permanent).
            glynn = GPUGlynnGrayPermanentCalculator(self._matrix)
            res = glynn.compute_permanent_batch([p[0] for p in pairs], [p[1] for p in pairs])
            for i, (s, t) in enumerate(pairs):
                if s.sum() == 0 and t.sum() != 0:
                    res[i] = 0.0
        return res


class GPURyserPermanentCalculator(_GuanGPUCalculator):
    """Ryser's formula over Guan codes with mode multiplicities
    (ryser_permanent_calculator.py:45-64 on bs_permanent_calculator_base.py:166-209)."""

    formula = L.TB_FORMULA_RYSER


class GPUChinHuhPermanentCalculator(_GuanGPUCalculator):
    """Chin-Huh (Glynn with multiplicities) over Guan codes
    (chin_huh_permanent_calculator.py:38-59)."""

    formula = L.TB_FORMULA_CHIN_HUH


class GPUClassicPermanentCalculator(BSPermanentCalculatorBase):
    """Counterpart of ClassicPermanentCalculator (classic_permanent_calculator.py:27-41):
    same edge values (no input particles -> 1 if no output particles else 0); the
    permanent itself comes from the GPU Ryser kernel rather than row expansion."""

    def compute_permanent(self) -> complex128:
        s, t = self._checked_states()
        if s.sum() == 0:
            return complex128(1 if t.sum() == 0 else 0)
        return GPURyserPermanentCalculator(self._matrix, s, t).compute_permanent()

    def compute_permanent_batch(self, input_states, output_states) -> np.ndarray:
        return GPURyserPermanentCalculator(self._matrix).compute_permanent_batch(
            input_states, output_states)
