"""GCC-A in the mode-assignment description.

Restates theboss's ModeAssignmentGeneralizedCliffordsSimulationStrategyInterface
(theboss/simulation_strategies/mode_assignment_generalized_cliffords_simulation_strategy.py:16-112):
the sample prefix is the ordered tuple of the modes drawn so far, the pmfs are memoised by
that tuple, the candidates of a layer are prefix + (i,) for every mode i (:98-112), and the
samples are returned as mode occupations (:41-51). A layer's pmf depends only on the
prefix's occupation, so prefixes that are permutations of each other share one computed
layer here; samples advance in lock-step as in GCC-A.
"""
from typing import List, Sequence, Tuple

import numpy as np

from ..states import mode_assignment_to_mode_occupation
from .generalized_cliffords_simulation_strategy import GeneralizedCliffordsSimulationStrategy


class ModeAssignmentGeneralizedCliffordsSimulationStrategyInterface(GeneralizedCliffordsSimulationStrategy):
    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        self._initialize_simulation(input_state)
        S, m = max(int(samples_number), 0), len(self.input_state)
        assignments: List[Tuple[int, ...]] = [tuple()] * S
        for _ in range(self.number_of_input_photons if S else 0):
            missing = sorted({a for a in assignments if a not in self.pmfs})
            by_occupation = {}
            for a in missing:
                by_occupation.setdefault(tuple(mode_assignment_to_mode_occupation(a, m)), []).append(a)
            if by_occupation:
                occ = sorted(by_occupation)
                _, pmfs = self._layer_pmfs(occ)
                for o, pmf in zip(occ, pmfs):
                    for a in by_occupation[o]:
                        self.pmfs[a] = pmf
                        self.possible_outputs[a] = [a + (i,) for i in range(m)]
            idx = self._draw(np.stack([self.pmfs[a] for a in assignments]))
            assignments = [a + (int(i),) for a, i in zip(assignments, idx)]
        if S:
            self.r_sample = assignments[-1]
            self.current_key = assignments[-1]
        return [tuple(mode_assignment_to_mode_occupation(a, m)) for a in assignments]
