"""DAC I/O (the CANS output format of cansio/dacdef*.f): files the reference
ships (the UTRIP spicule run's t.dac and x.dac, copied as data fixtures) read
and re-write byte for byte, and round trips of every shape and type."""
from pathlib import Path

import numpy as np
import pytest

from utrip_amd import dac

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_reads_reference_time_series():
    snaps, h = dac.read(GOLDEN / "spicule_t.dac")
    assert h == {"version": 0, "mtype": 6, "dims": [-1], "series": True}
    assert [float(s) for s in snaps] == [0.0, 0.18685411175615266]


def test_reads_reference_static_grid():
    snaps, h = dac.read(GOLDEN / "spicule_x.dac")
    assert h["dims"] == [1026] and not h["series"] and len(snaps) == 1
    x = snaps[0]
    assert x.shape == (1026,) and np.all(np.diff(x) > 0)


@pytest.mark.parametrize("name", ["spicule_t.dac", "spicule_x.dac"])
def test_rewrite_is_byte_identical(tmp_path, name):
    snaps, h = dac.read(GOLDEN / name)
    out = tmp_path / name
    dac.write(out, snaps, series=h["series"])
    assert out.read_bytes() == (GOLDEN / name).read_bytes()


@pytest.mark.parametrize("shape,dtype", [((7,), np.float64), ((5, 3), np.float64), ((4, 3, 2), np.float32),
                                         ((6,), np.int32)])
def test_round_trip(tmp_path, shape, dtype):
    rng = np.random.default_rng(0)
    snaps = [np.asfortranarray((rng.standard_normal(shape) * 100).astype(dtype)) for _ in range(3)]
    p = tmp_path / "a.dac"
    dac.write(p, snaps)
    back, h = dac.read(p)
    assert h["dims"] == list(shape) + [-1]
    for a, b in zip(snaps, back):
        np.testing.assert_array_equal(a, b)
    assert np.array_equal(dac.read_last(p), snaps[-1])
