"""MI355X-native force-evaluation hot path of PotentialFrameWork.jl.

The product is ``libpfw_mi355x.so`` (HIP kernels for gfx950 behind the C ABI
in include/pfw_mi355x.h); this package is the host-side mirror of the
reference's ``Workspace`` / ``evaluate!`` / AtomsCalculators API over ctypes.

Because the directory name contains a dot, import it by path, e.g.::

    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "pfw_amd", "potentialframework.jl_amd/__init__.py",
        submodule_search_locations=["potentialframework.jl_amd"])
"""
from ._capi import (  # noqa: F401
    ArgumentError,
    CapacityError,
    DeviceError,
    DimensionMismatch,
    DomainError,
    StaleError,
    load as load_library,
    LIB_PATH,
)
from .system import System  # noqa: F401
from . import workloads  # noqa: F401
from .neighbourlist import Workspace, neighbour_list, neighbours, NeighbourSnapshot, pair_set_digest  # noqa: F401
from .pair import (  # noqa: F401
    SmoothPair,
    RadialPotential,
    LennardJonesRadial,
    lennard_jones,
    EnergyOnly,
    EnergyForcesVirial,
    evaluate_,
    evaluate,
    evaluate_batch,
)
from .ace import ACECalculator, ResolvedModel, load_resolved, ace_evaluate, project_species  # noqa: F401
from . import ace, ace_basis  # noqa: F401
from .adjoints import (  # noqa: F401
    pair_parameter_vjp,
    ace_parameter_vjp,
    rrule_pair_evaluate,
    rrule_ace_evaluate,
    fitting_loss,
)
from .fitting import design_matrix, training_rows, linear_parameters, install_readout  # noqa: F401
from .precision import compare_precision, PrecisionReport  # noqa: F401
from . import features  # noqa: F401
from . import tune  # noqa: F401
from .tune import autotune  # noqa: F401
from .features import feature_vjp, ace_features_rrule  # noqa: F401
from .features import ACEFeatureLayer, ace_features, feature_parameters, feature_basis_summary  # noqa: F401
from .calculators import (  # noqa: F401
    potential_energy,
    forces,
    virial,
    energy_forces,
    energy_forces_virial,
    forces_,
    energy_forces_,
    energy_forces_virial_,
    calculate,
)
