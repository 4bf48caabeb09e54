"""Ensemble sharding across devices (host logic only; nothing here runs on a GPU): every
instance goes to exactly one device, the assignment is deterministic, and the loads obey the
longest-processing-time bound."""
import pytest

from neofiz_amd import ensemble


@pytest.mark.parametrize("devices", [1, 2, 3, 8])
def test_every_instance_once_and_balanced(devices):
    inst = ensemble.draw(256, seed=2024)
    parts = ensemble.shard(inst, devices)
    assert len(parts) == devices
    flat = sorted(i for p in parts for i in p)
    assert flat == list(range(len(inst)))
    assert all(p == sorted(p) for p in parts)
    cost = [sum(ensemble.predicted_cost(inst[i]) for i in p) for p in parts]
    biggest = max(ensemble.predicted_cost(x) for x in inst)
    mean = sum(cost) / devices
    assert max(cost) <= mean + biggest
    assert ensemble.shard(inst, devices) == parts


def test_cost_follows_mesh_refinement():
    small = ensemble.Instance(8, 80, 190.0, 0.1, (1.0,) * 5)
    large = ensemble.Instance(16, 160, 190.0, 0.1, (1.0,) * 5)
    assert ensemble.predicted_cost(large) == 8 * ensemble.predicted_cost(small)


def test_one_instance_per_device_and_errors():
    inst = ensemble.draw(3, seed=1)
    parts = ensemble.shard(inst, 5)
    assert sorted(len(p) for p in parts) == [0, 0, 1, 1, 1]
    with pytest.raises(ValueError):
        ensemble.shard(inst, 0)
