"""Pins oracle/dyn.py (GPMA+ update semantics) on hand-computed sequences:
insert, weight change, delete of present and absent edges, the last
occurrence of a key winning inside one batch, weight-0 edges at creation,
and the CSR in both directions."""
import numpy as np

from oracle.dyn import EdgeDict


def test_insert_modify_delete_sequence():
    d = EdgeDict(4)
    c = d.apply([0, 0, 2, 3], [1, 2, 3, 3], [0.5, 0.25, 1.0, 2.0])  # a self-loop is an ordinary edge
    assert c == {"inserted": 4, "deleted": 0, "modified": 0, "distinct": 4}
    c = d.apply([0, 1, 2], [1, 0, 3], [0.75, 0.0, 0.0])  # change, delete absent (no-op), delete
    assert c == {"inserted": 0, "deleted": 1, "modified": 1, "distinct": 3}
    assert d.e == {(0, 1): 0.75, (0, 2): 0.25, (3, 3): 2.0}
    rp, col, w = d.csr()
    np.testing.assert_array_equal(rp, [0, 2, 2, 2, 3])
    np.testing.assert_array_equal(col, [1, 2, 3])
    np.testing.assert_array_equal(w, [0.75, 0.25, 2.0])
    rp, col, w = d.csr(reverse=True)  # in-rows: sources of each target
    np.testing.assert_array_equal(rp, [0, 0, 1, 2, 3])
    np.testing.assert_array_equal(col, [0, 0, 3])
    np.testing.assert_array_equal(w, [0.75, 0.25, 2.0])


def test_last_occurrence_in_a_batch_wins():
    d = EdgeDict(3)
    d.apply([0, 0, 0], [1, 1, 1], [1.0, 0.0, 3.0])  # insert, delete, insert: ends present with 3.0
    assert d.e == {(0, 1): 3.0}
    c = d.apply([0, 0], [1, 1], [5.0, 0.0])  # ends deleted
    assert c["deleted"] == 1 and d.e == {}


def test_creation_drops_zero_weights_and_defaults_to_one():
    d = EdgeDict(3, [0, 2, 3, 3], [1, 2, 0], [0.0, 0.5, 2.0])
    assert d.e == {(0, 2): 0.5, (1, 0): 2.0}
    d = EdgeDict(2, [0, 1, 1], [1], None)
    assert d.e == {(0, 1): 1.0}
    d.apply([1], [0])  # no weights: 1.0
    assert d.e == {(0, 1): 1.0, (1, 0): 1.0}
