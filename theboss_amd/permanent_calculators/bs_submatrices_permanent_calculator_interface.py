"""Interface of submatrices permanents calculators.

Mirrors theboss's BSSubmatricesPermanentCalculatorInterface
(theboss/boson_sampling_utilities/permanent_calculators/bs_submatrices_permanent_calculator_interface.py:12-50):
compute_permanents() returns, for every input mode i, the permanent of the effective
matrix with one particle removed from input mode i (one value per mode).
"""
import abc
from typing import List

from numpy import complex128, ndarray


class BSSubmatricesPermanentCalculatorInterface(abc.ABC):
    @abc.abstractmethod
    def compute_permanents(self) -> List[complex128]:
        """Permanents of the input-particle-deleted effective submatrices."""

    @property
    @abc.abstractmethod
    def matrix(self) -> ndarray:
        ...

    @matrix.setter
    @abc.abstractmethod
    def matrix(self, matrix: ndarray) -> None:
        ...

    @property
    @abc.abstractmethod
    def input_state(self) -> ndarray:
        ...

    @input_state.setter
    @abc.abstractmethod
    def input_state(self, input_state: ndarray) -> None:
        ...

    @property
    @abc.abstractmethod
    def output_state(self) -> ndarray:
        ...

    @output_state.setter
    @abc.abstractmethod
    def output_state(self, output_state: ndarray) -> None:
        ...
