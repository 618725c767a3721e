"""The measurement entry points the bench reports beside the vendor peak:
pearl_stream_probe times a plain device copy, pearl_stream_mix a bare kernel
with a cell kernel's read and write planes.  A 1 GiB copy does not fit in
the 256 MB last-level cache, so its rate is HBM-bound and must lie below
MI355X's 8 TB/s peak; a copy kernel that ran at all moves well over 1 TB/s.
The same holds for each kernel's planes at the production size (805 MB for
prepare's twelve).  A device that does not exist is an error, not a fault,
and leaves no error behind for the next entry point."""
import ctypes

import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def test_stream_probe_rate_is_hbm_bound(pkg):
    gbps = pkg.stream_probe(0, 1 << 30, 5)
    print("copy ceiling", gbps, "GB/s", flush=True)
    assert 1000.0 < gbps < 8000.0


def test_stream_probe_rejects_missing_device(pkg):
    g = ctypes.c_double(1.0)
    assert pkg.lib().pearl_stream_probe(4096, 1 << 20, 1, ctypes.byref(g)) == pkg.ERR_DEVICE
    assert g.value == 0.0
    # the next entry point's launch checks see no stale error
    gd = load_golden("tiny_3x2")
    with pkg.Solver(gd["pd"], float(gd["sc0"][9])) as s:
        s.set_state(gd["phi0"], gd["c0"], gd["mu0"], gd["sc0"][:9])
        info = s.advance(float(gd["end_time"]))
    assert info.accepted_dt > 0


@pytest.mark.parametrize("reads,writes", [(4, 3), (5, 7), (6, 2)])
def test_stream_mix_rate_is_hbm_bound(pkg, reads, writes):
    us, gbps = pkg.stream_mix(reads, writes, 4096, 2048, 4, 5)
    print("bare pattern", reads, writes, us, "us", gbps, "GB/s", flush=True)
    assert 1000.0 < gbps < 8000.0
    assert us == pytest.approx((reads + writes) * 8 * 4096 * 2048 / (gbps * 1e3), rel=1e-9)


def test_stream_mix_ragged_sizes_and_missing_device(pkg):
    # a width off the tile multiple and a height off the row multiple
    us, gbps = pkg.stream_mix(2, 2, 1000, 333, 8, 2)
    assert us > 0 and gbps > 0
    u, g = ctypes.c_double(1.0), ctypes.c_double(1.0)
    assert pkg.lib().pearl_stream_mix(4096, 1, 1, 64, 8, 4, 1, ctypes.byref(u), ctypes.byref(g)) == pkg.ERR_DEVICE
    assert u.value == 0.0 and g.value == 0.0
