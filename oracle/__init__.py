"""The oracle: CPU restatements of PotentialFrameWork.jl used to check the HIP path.

This package is test infrastructure. Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may load it, and only as the checker
or the timed CPU baseline; the product library never links or calls it.

Contents
--------
* ``liboracle.so`` (``nl_oracle.cpp``, ``pair_oracle.cpp``, ``ace_oracle.cpp``): FP64
  C++ restatements of the reference CPU path:
  - ``prepare!`` / ``image_topology!`` / ``minimum_topology!`` / ``schedule!``
    (src/submodules/NeighbourList/storage.jl:83-306),
  - ``visit_directed_images!`` (src/submodules/NeighbourList/interface.jl:116-143),
  - ``cell_task!`` / ``execute!`` (src/submodules/NeighbourList/execution.jl:169-238),
  - SmoothPair and energy-shifted Lennard-Jones callbacks
    (src/submodules/PairPotentials/models.jl:153-167, lennard_jones.jl:5-17),
  - the ACE site pipeline (see ``ace.py``).
* ``brute.py``: independent O(N^2) and exhaustive-image oracles restated from the
  reference's own tests (test/fixtures.jl:28-51, test/periodic_images.jl:3-28).

Pinning: every function here is checked in ``tests/test_oracle_*.py`` against the
reference's golden vectors, closed forms or brute force (see the module docstrings).
"""
from .core import (  # noqa: F401
    lib,
    nl_pairs,
    pair_evaluate,
    KIND_SMOOTH,
    KIND_LJ,
    OracleError,
)
