"""Pins of the CPU restatement of ReadLisFile (src/tree.hpp:1540-1583):
round trips through the synthetic .lis writer (an independent encoder that
assigns (cpu_id, id_in_run) so that the sorted order is the snapshot's), a
hand-made file whose order, key cuts and header fields are worked out by
hand, and the short-file error."""
import os

import numpy as np
import pytest

import oracle_py as O
from plan_amd import synth


@pytest.fixture(scope="module")
def snap():
    return synth.small_case(3000, 3, seed=2)


@pytest.mark.parametrize("num_files,high_bits", [(1, False), (1, True), (3, False), (5, True)])
def test_round_trip(tmp_path, snap, num_files, high_bits):
    paths = synth.write_lis(snap, str(tmp_path / "snap"), num_files=num_files, seed=7, high_bits=high_bits)
    assert len(paths) == num_files
    r = O.read_lis(paths)
    assert np.array_equal(r["pos"], snap.pos.astype(np.float64))
    assert np.array_equal(r["vel"], snap.vel.astype(np.float64))
    assert np.array_equal(r["density"], snap.density.astype(np.float64))
    assert np.array_equal(r["property_index"], snap.property_index)
    assert np.array_equal(r["id"], np.arange(snap.n, dtype=np.uint32))
    assert list(r["box_min"]) == snap.params["box_min"] and list(r["box_max"]) == snap.params["box_max"]
    assert r["dt"] == float(np.float32(1e-3)) and r["num_types"] == 1


def _hand_file(path, n_declared=None):
    """Five records; grid limits differ from the domain limits, two species."""
    grid = [-1, 1, -2, 2, -3, 3]
    dom = [-0.5, 0.25, -0.75, 0.5, -0.125, 0.375]
    keys = [(1, 5), (0, 7), (0, 2), (65536, 3), (1, (1 << 32) + 1)]
    rec = np.zeros(5, synth.LIS_RECORD)
    for i, (c, r) in enumerate(keys):
        rec[i]["pos"] = [i + 0.5, -i, 2 * i]
        rec[i]["vel"] = [10 * i, 0.25, -i]
        rec[i]["density"] = 100 + i
        rec[i]["property_index"] = i % 2
        rec[i]["id_in_run"] = r
        rec[i]["cpu_id"] = c
    head = (np.array(grid + dom, "<f4").tobytes() + np.array([2], "<i4").tobytes()
            + np.array([3.0, 1.0], "<f4").tobytes() + np.array([0.625, 0.01], "<f4").tobytes()
            + np.array([5 if n_declared is None else n_declared], "<i8").tobytes())
    with open(path, "wb") as f:
        f.write(head + rec.tobytes())


def test_hand_made_file(tmp_path):
    p = str(tmp_path / "hand.lis")
    _hand_file(p)
    assert os.path.getsize(p) == 48 + 4 + 8 + 8 + 8 + 5 * 44
    r = O.read_lis(p)
    # keys after the cuts: cpu_id to 16 bits, id_in_run to 32 bits:
    # (1,5) (0,7) (0,2) (0,3) (1,1)  ->  sorted (0,2) (0,3) (0,7) (1,1) (1,5)
    want = [2, 3, 1, 4, 0]
    assert np.array_equal(r["pos"][:, 0], np.array(want) + 0.5)
    assert np.array_equal(r["density"], 100.0 + np.array(want))
    assert np.array_equal(r["property_index"], np.array(want) % 2)
    assert np.array_equal(r["id"], np.arange(5))
    assert list(r["box_min"]) == [-0.5, -0.75, -0.125]
    assert list(r["box_max"]) == [0.25, 0.5, 0.375]
    assert list(r["type_info"]) == [3.0, 1.0]
    assert r["time"] == 0.625 and r["dt"] == float(np.float32(0.01))


def test_short_file_is_an_error(tmp_path):
    p = str(tmp_path / "short.lis")
    _hand_file(p, n_declared=6)
    with pytest.raises(ValueError):
        O.read_lis(p)
