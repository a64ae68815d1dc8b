"""GPU Clifford & Clifford submatrices calculator.

Counterpart of theboss's BSCCRyserSubmatricesPermanentCalculator
(theboss/boson_sampling_utilities/permanent_calculators/bs_cc_ryser_submatrices_permanent_calculator.py:16-119)
on the base semantics of bs_submatrices_permanent_calculator_base.py:17-201:
compute_permanents() returns one complex128 per input mode i, the permanent of the
effective matrix (rows = output particles, columns = input particles) with one particle
removed from mode i, and 0 for unoccupied modes. A single input particle returns the
input state itself (:156-158).
"""
from typing import List, Optional, Sequence

import numpy as np
from numpy import complex128

from .. import _device as D
from .. import ops
from ..states import as_occupation, compact
from .bs_submatrices_permanent_calculator_interface import BSSubmatricesPermanentCalculatorInterface


class BSSubmatricesPermanentCalculatorBase(BSSubmatricesPermanentCalculatorInterface):
    """bs_submatrices_permanent_calculator_base.py:17-65: stores references, None -> []."""

    def __init__(self, matrix, input_state: Optional[Sequence[int]] = None,
                 output_state: Optional[Sequence[int]] = None) -> None:
        if output_state is None:
            output_state = []
        if input_state is None:
            input_state = []
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


class GPUBSCCRyserSubmatricesPermanentCalculator(BSSubmatricesPermanentCalculatorBase):
    """All input-particle-deleted minors in one GPU sweep (binary Gray sweep for
    collision-free inputs, mixed-radix Guan sweep otherwise)."""

    def compute_permanents(self) -> List[complex128]:
        s = as_occupation(self._input_state)
        if int(s.sum()) == 1:
            return [complex128(v) for v in self._input_state]
        return [complex128(v) for v in self.minors_array()]

    def minors_array(self) -> np.ndarray:
        """compute_permanents as one complex128 array of length m (no per-entry list)."""
        s = as_occupation(self._input_state)
        if int(s.sum()) == 1:
            return s.astype(np.complex128)
        return self._minors([(s, as_occupation(self._output_state))])[0]

    def compute_permanents_batch(self, input_states, output_states) -> np.ndarray:
        """Minors for many (input, output) pairs on the same matrix; all pairs must have
        the same particle numbers. Returns len(pairs) x m."""
        return self._minors([(as_occupation(s), as_occupation(t))
                             for s, t in zip(input_states, output_states)])

    def _minors(self, pairs) -> np.ndarray:
        U = np.asarray(self._matrix, dtype=np.complex128)
        B = len(pairs)
        m = len(pairs[0][0])
        ks = {int(p[0].sum()) for p in pairs}
        kt = {int(p[1].sum()) for p in pairs}
        if len(ks) != 1 or len(kt) != 1:
            raise ValueError("all items of a batch need the same particle numbers")
        k, kout = ks.pop(), kt.pop()
        out = np.zeros((B, m), dtype=np.complex128)
        if k == 0:
            return out
        if k == 1:
            for b, (s, _) in enumerate(pairs):
                out[b] = s
            return out
        comps = [(compact(t), compact(s)) for s, t in pairs]
        J = max(1, max(len(c[0][0]) for c in comps))
        Dd = max(1, max(len(c[1][0]) for c in comps))
        collision_free = all(np.all(c[1][1] == 1) for c in comps) and kout == k - 1
        stack = np.zeros((B * J, Dd), dtype=np.complex128)
        out_modes = np.zeros((B, J), dtype=np.int32)
        out_mult = np.zeros((B, J), dtype=np.int32)
        in_modes = np.zeros((B, Dd), dtype=np.int32)
        in_mult = np.zeros((B, Dd), dtype=np.int32)
        for b, ((om, ot), (im, it)) in enumerate(comps):
            stack[b * J:b * J + len(om), :len(im)] = U[np.ix_(om, im)]
            out_modes[b] = b * J + np.arange(J)
            out_mult[b, :len(ot)] = ot
            in_modes[b] = np.arange(Dd)
            in_mult[b, :len(it)] = it
        # the gathered matrix block and the four index arrays in one upload (each pageable
        # upload waits for the stream): bytes, viewed as complex128 and int32 on the device
        T = D.require_gpu()
        ints = np.concatenate([out_modes.ravel(), out_mult.ravel(), in_modes.ravel(),
                               in_mult.ravel()])
        buf = T.from_numpy(np.concatenate([stack.view(np.uint8).ravel(),
                                           ints.view(np.uint8)])).to(D.device())
        cb = stack.nbytes  # a multiple of 16: the int32 view starts aligned
        Vd = buf[:cb].view(T.complex128).view(B * J, Dd)
        ints_d = buf[cb:].view(T.int32)
        oj, dj = B * J, B * Dd
        minors = ops.cc_submatrices_batch(
            Vd, ints_d[:oj].view(B, J), ints_d[oj:2 * oj].view(B, J),
            ints_d[2 * oj:2 * oj + dj].view(B, Dd), ints_d[2 * oj + dj:].view(B, Dd), k,
            collision_free).cpu().numpy()
        for b, (_, (im, _)) in enumerate(comps):
            out[b, im] = minors[b, :len(im)]
        return out
