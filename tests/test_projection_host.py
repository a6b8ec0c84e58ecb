"""Host logic of project_species (src/submodules/ACE/interface.jl:115-121):
argument checks with the reference's error classes (test/ace_interface.jl:76-78),
the channels and orders it records, and the policy it writes for checkpoints
(interface_persistence.jl:3-6). No device calls."""
import os

import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def tial(pfw):
    return pfw.load_resolved(os.path.join(GOLD, "resolved_model_v1.toml"))


def test_project_species_records_channels_and_orders(pfw, tial):
    m, ps, _ = tial
    ti, al = m.atomic_numbers.index(22), m.atomic_numbers.index(13)
    p = pfw.project_species(m, ["Al", "Ti", "Ti"], range(1, 4))
    assert p.species == tuple(sorted((ti, al))) and p.orders == (1, 3)
    assert m.species == () and m.orders is None  # the original is not changed
    assert pfw.project_species(m, 22, range(0, 1)).species == (ti,)
    assert pfw.project_species(m, "Al", range(2, 3)).orders == (2, 2)
    assert pfw.project_species(m, [], range(0, 4)).species == ()
    pol = p.source["interface_policy"]
    assert pol["species"] == [c + 1 for c in p.species] and pol["orders"] == [1, 3]
    assert p.order_allowed(1) and not p.order_allowed(0) and m.order_allowed(0)


def test_project_species_errors(pfw, tial):
    m, _, _ = tial
    with pytest.raises(pfw.ArgumentError):
        pfw.project_species(m, ["H"], range(0, 2))  # no such element channel
    with pytest.raises(pfw.ArgumentError):
        pfw.project_species(m, ["Ti"], range(-1, 3))  # negative order
    with pytest.raises(pfw.ArgumentError):
        pfw.project_species(m, ["Ti"], [0, 1])  # not a unit range
    with pytest.raises(pfw.ArgumentError):
        pfw.project_species(m, ["Ti"], range(0, 4, 2))
    with pytest.raises(pfw.ArgumentError):
        pfw.project_species(m, ["Ti"], range(2, 2))  # empty
    with pytest.raises(pfw.ArgumentError):
        pfw.project_species(m, ["Ti-48"], range(0, 1))  # isotopes/custom names are not element channels
