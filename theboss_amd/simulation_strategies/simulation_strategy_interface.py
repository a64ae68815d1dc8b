"""theboss/simulation_strategies/simulation_strategy_interface.py:7-23."""
import abc
from typing import List, Sequence, Tuple


class SimulationStrategyInterface(abc.ABC):
    @abc.abstractmethod
    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        """Samples (output occupation tuples) of the experiment with this input state."""
