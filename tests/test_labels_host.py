"""Species-label mapping on the host (bind_species, src/submodules/PairPotentials/
models.jl:50-58; element channels, src/submodules/ACE/environments.jl:1-10):
the cached distinct-label codes of a System, channels by a model's species
list, first-appearance order for single-species models, and the error
classes for unknown, isotope-like and unhashable labels."""
import numpy as np
import pytest


def test_label_codes_and_channels(pfw):
    from pfw_amd.system import Labels, channels, map_labels, unique_in_order
    labels = ["B", "A", "B", "C", "A"]
    s = pfw.System(np.zeros((5, 3)), 5 * np.eye(3), (1, 1, 1), labels)
    assert isinstance(s.labels, Labels) and list(s.labels) == labels
    distinct, codes = s.labels.codes()
    assert distinct == ["B", "A", "C"] and codes.tolist() == [0, 1, 0, 2, 1]
    assert s.labels.codes()[1] is codes  # computed once
    assert channels(s.labels, ["A", "B", "C"]).tolist() == [1, 0, 1, 2, 0]
    assert channels(labels, ["C", "B", "A"]).tolist() == [1, 2, 1, 0, 2]  # plain lists too
    assert unique_in_order(labels) == ["B", "A", "C"]
    with pytest.raises(pfw.ArgumentError, match="C"):
        channels(s.labels, ["A", "B"])
    with pytest.raises(pfw.ArgumentError):
        map_labels([["not", "hashable"]], lambda x: 0)
    empty = pfw.System(np.zeros((0, 3)), 5 * np.eye(3), (1, 1, 1), [])
    assert channels(empty.labels, ["A"]).shape == (0,)


def test_ace_and_feature_channels(pfw):
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    model, _, _ = pfw.load_resolved(os.path.join(gold, "resolved_model_v1.toml"))
    labels = ["Ti", "Al", "Ti", "Al", "Al"]
    want = [model.atomic_numbers.index(22 if s == "Ti" else 13) for s in labels]
    assert model.channels(labels).tolist() == want
    with pytest.raises(pfw.ArgumentError):
        model.channels(["Ti", "Si"])  # an element without a channel
    with pytest.raises(pfw.ArgumentError):
        model.channels(["Ti", "Ti2"])  # not an element label
    F = pfw.features
    layer = F.ACEFeatureLayer(["Si", "O"], F.SmoothExponential([(1, 0), (1, 1)], 4.0, [0, 1]), 2, 1,
                              [dict(L=0, parity="even", channels=1)])
    s = pfw.System(np.zeros((3, 3)), 5 * np.eye(3), (1, 1, 1), ["O", "Si", "O"])
    assert F._channels(layer, s).tolist() == [1, 0, 1]
    with pytest.raises(pfw.ArgumentError):
        F._channels(layer, pfw.System(np.zeros((1, 3)), 5 * np.eye(3), (1, 1, 1), ["Ti"]))
