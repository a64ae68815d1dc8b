"""Permanent calculators with theboss's interfaces, computed by gfx950 kernels."""
from .bs_permanent_calculator_factory import BSPermanentCalculatorFactory, PermanentCalculatorType
from .bs_permanent_calculator_interface import BSPermanentCalculatorInterface
from .bs_submatrices_permanent_calculator_interface import BSSubmatricesPermanentCalculatorInterface
from .gpu_permanent_calculators import (
    BSPermanentCalculatorBase,
    GPUChinHuhPermanentCalculator,
    GPUClassicPermanentCalculator,
    GPUGlynnGrayFP32PermanentCalculator,
    GPUGlynnGrayPermanentCalculator,
    GPURyserPermanentCalculator,
)
from .gpu_submatrices_calculators import (
    BSSubmatricesPermanentCalculatorBase,
    GPUBSCCRyserSubmatricesPermanentCalculator,
)

# Reference class names resolve to the GPU implementations, so code written against
# theboss's calculators can import these names unchanged.
GlynnGrayPermanentCalculator = GPUGlynnGrayPermanentCalculator
RyserPermanentCalculator = GPURyserPermanentCalculator
ChinHuhPermanentCalculator = GPUChinHuhPermanentCalculator
ClassicPermanentCalculator = GPUClassicPermanentCalculator
BSCCRyserSubmatricesPermanentCalculator = GPUBSCCRyserSubmatricesPermanentCalculator
