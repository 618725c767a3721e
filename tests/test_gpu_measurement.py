"""The measurement entry point the bench reports beside the vendor peak:
pearl_stream_probe times a plain device copy.  A 1 GiB copy does not fit in
the 256 MB last-level cache, so its rate is HBM-bound and must lie below
MI355X's 8 TB/s peak; a copy kernel that ran at all moves well over 1 TB/s.
A device that does not exist is an error, not a fault, and leaves no error
behind for the next entry point."""
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
