"""Sampling strategies built on the GPU permanent calculators."""
from .batched_gccb_simulation_strategy import BatchedGCCBSimulationStrategy
from .bobs_strategies import (
    LossyStateApproximationSimulationStrategy,
    NonuniformLossesApproximationStrategy,
)
from .generalized_cliffords_b_simulation_strategy import GeneralizedCliffordsBSimulationStrategy
from .generalized_cliffords_simulation_strategy import GeneralizedCliffordsSimulationStrategy
from .generalized_cliffords_uniform_losses_simulation_strategy import (
    GeneralizedCliffordsUniformLossesSimulationStrategy,
)
from .lossy_strategies import (
    GeneralizedCliffordsBUniformLossesSimulationStrategy,
    LossyNetworksGeneralizedCliffordsSimulationStrategy,
)
from .mode_assignment_generalized_cliffords_simulation_strategy import (
    ModeAssignmentGeneralizedCliffordsSimulationStrategyInterface,
)
from .simulation_strategy_factory import SimulationStrategyFactory, StrategyType
from .simulation_strategy_interface import SimulationStrategyInterface
