"""Interface of boson-sampling permanent calculators.

Mirrors theboss's BSPermanentCalculatorInterface
(theboss/boson_sampling_utilities/permanent_calculators/bs_permanent_calculator_interface.py:18-52):
a calculator holds an interferometer matrix and input/output occupation vectors and
computes the permanent of the effective scattering matrix they select. Setting both
states to all ones gives the permanent of the matrix itself.
"""
import abc
from typing import Sequence


class BSPermanentCalculatorInterface(abc.ABC):
    @abc.abstractmethod
    def compute_permanent(self) -> complex:
        """Permanent of the effective scattering matrix of (matrix, input, output)."""

    @property
    @abc.abstractmethod
    def matrix(self) -> Sequence[Sequence[complex]]:
        ...

    @matrix.setter
    @abc.abstractmethod
    def matrix(self, matrix: Sequence[Sequence[complex]]) -> None:
        ...

    @property
    @abc.abstractmethod
    def input_state(self) -> Sequence[int]:
        ...

    @input_state.setter
    @abc.abstractmethod
    def input_state(self, input_state: Sequence[int]) -> None:
        ...

    @property
    @abc.abstractmethod
    def output_state(self) -> Sequence[int]:
        ...

    @output_state.setter
    @abc.abstractmethod
    def output_state(self, output_state: Sequence[int]) -> None:
        ...
