"""The brick pool policy restated in numpy (oracle BrickPoolOracle), pinned by
hand-derived cases of its definition: first appearance order, lowest free
slots, reuse of freed slots, all-or-nothing failure, skipped releases."""
import numpy as np
import pytest


def fresh(oracle_mod, g=4, brick=2, cap=6):
    table = np.full(g ** 3, -1, np.int32)
    params = np.zeros((cap * brick ** 3, 4), np.float32)
    return oracle_mod.BrickPoolOracle(table, params, brick)


def test_first_appearance_and_lowest_slots(oracle_mod):
    p = fresh(oracle_mod)
    slots, new = p.alloc([5, 3, 5, 9])
    assert new == 3 and slots.tolist() == [0, 1, 0, 2]
    assert p.release([3, 3, 40]) == 1  # duplicates and unallocated bricks are skipped
    slots, new = p.alloc([7, 8, 5])
    assert new == 2 and slots.tolist() == [1, 3, 0]  # the freed slot 1 first, then 3


def test_failure_changes_nothing(oracle_mod):
    p = fresh(oracle_mod, cap=3)
    p.alloc([1, 2])
    before = p.table.copy()
    with pytest.raises(ValueError):
        p.alloc([10, 11])  # two new bricks, one free slot
    assert np.array_equal(p.table, before)
    with pytest.raises(ValueError):
        p.alloc([64])  # out of range for 4^3 bricks


def test_update_targets_slots(oracle_mod):
    p = fresh(oracle_mod)
    p.alloc([6, 2])
    packed = np.arange(3 * 8 * 4, dtype=np.float32).reshape(3 * 8, 4)
    p.update([2, 6, 33], packed)  # 33 has no slot: skipped
    assert np.array_equal(p.params[8:16], packed[0:8])   # brick 2 -> slot 1
    assert np.array_equal(p.params[0:8], packed[8:16])   # brick 6 -> slot 0
