"""Drop-in mirror of ``ionization.mesh`` whose evolve step runs on the GPU."""

from .meshes import QuantumMesh, LineMesh, SphericalHarmonicMesh, CylindricalSliceMesh, WrappingDirection
from .mesh_operators import (
    MeshOperators,
    KineticEnergyDerivation,
    LineLengthGaugeOperators,
    LineVelocityGaugeOperators,
    CylindricalSliceLengthGaugeOperators,
    SphericalHarmonicLengthGaugeOperators,
    SphericalHarmonicVelocityGaugeOperators,
)
from .evolution_methods import EvolutionMethod, SplitInteractionOperator, AlternatingDirectionImplicit
from .sims import (
    MeshSimulation,
    MeshSpecification,
    LineSpecification,
    CylindricalSliceSpecification,
    SphericalHarmonicSimulation,
    SphericalHarmonicSpecification,
)
from .data import Data, Datastore, Fields, Norm, InnerProducts, NormBySphericalHarmonic, DEFAULT_DATASTORE_TYPES
