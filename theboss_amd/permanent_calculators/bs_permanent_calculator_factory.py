"""Calculator factory.

Mirrors theboss's BSPermanentCalculatorFactory and PermanentCalculatorType
(theboss/boson_sampling_utilities/permanent_calculators/bs_permanent_calculator_factory.py:29-115):
same enum values (CLASSIC=1, GLYNN=2, CHIN_HUH=3, RYSER=4), default RYSER, unknown types
fall back to Chin-Huh (:84-87). Every type builds a GPU calculator; GLYNN_FP32 is an
added type for the FP32 column-sum Glynn variant.
"""
import enum
from typing import Optional

from numpy import ndarray

from .bs_permanent_calculator_interface import BSPermanentCalculatorInterface
from .gpu_permanent_calculators import (
    GPUChinHuhPermanentCalculator,
    GPUClassicPermanentCalculator,
    GPUGlynnGrayFP32PermanentCalculator,
    GPUGlynnGrayPermanentCalculator,
    GPURyserPermanentCalculator,
)


class PermanentCalculatorType(enum.IntEnum):
    CLASSIC = enum.auto()
    GLYNN = enum.auto()
    CHIN_HUH = enum.auto()
    RYSER = enum.auto()
    GLYNN_FP32 = enum.auto()


class BSPermanentCalculatorFactory:
    def __init__(self, matrix: Optional[ndarray], input_state: Optional[ndarray],
                 output_state: Optional[ndarray],
                 calculator_type: PermanentCalculatorType = PermanentCalculatorType.RYSER):
        self._matrix = matrix
        self._input_state = input_state
        self._output_state = output_state
        self._calculator_type = calculator_type
        self._calculator_mapping = {
            PermanentCalculatorType.CLASSIC: GPUClassicPermanentCalculator,
            PermanentCalculatorType.GLYNN: GPUGlynnGrayPermanentCalculator,
            PermanentCalculatorType.CHIN_HUH: GPUChinHuhPermanentCalculator,
            PermanentCalculatorType.RYSER: GPURyserPermanentCalculator,
            PermanentCalculatorType.GLYNN_FP32: GPUGlynnGrayFP32PermanentCalculator,
        }

    @property
    def matrix(self) -> ndarray:
        return self._matrix

    @matrix.setter
    def matrix(self, new_matrix: ndarray) -> None:
        self._matrix = new_matrix

    @property
    def input_state(self) -> ndarray:
        return self._input_state

    @input_state.setter
    def input_state(self, new_input_state: ndarray) -> None:
        self._input_state = new_input_state

    @property
    def output_state(self) -> ndarray:
        return self._output_state

    @output_state.setter
    def output_state(self, new_output_state: ndarray) -> None:
        self._output_state = new_output_state

    def generate_calculator(self) -> BSPermanentCalculatorInterface:
        cls = self._calculator_mapping.get(self._calculator_type, GPUChinHuhPermanentCalculator)
        return cls(matrix=self.matrix, input_state=self.input_state, output_state=self.output_state)
