"""pearl_diagnostics_noalloc (no device allocation; its scratch is a static
device array) against pearl_diagnostics on the same states: every field
bitwise equal (the same kernels and blocked summation order), on two calls
in a row, the second on a smaller domain, at two tile heights; an oversized
domain is rejected before any device work; a state with a non-finite phase
fraction reports the invalid mixture like the context's call."""
import numpy as np
import pytest
import torch

from conftest import load_golden

pytestmark = pytest.mark.gpu


def device_state(g):
    t = [torch.from_numpy(np.ascontiguousarray(g[k], dtype=np.float64)).cuda() for k in ("phi0", "c0", "mu0")]
    torch.cuda.synchronize()
    return t


def context_diagnostics(pkg, g, k):
    with pkg.Solver(g["pd"], float(g["sc0"][9]), cells_per_thread=k) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        return s.diagnostics()


@pytest.mark.parametrize("k", [4, 1])
def test_noalloc_diagnostics_equal_context_diagnostics(pkg, k):
    # the larger domain first, then a smaller one: the second call's
    # partials and totals must not pick up anything the first left behind
    for name in ("colony_256x128", "small_48x40", "tiny_3x2"):
        g = load_golden(name)
        print("no-alloc diagnostics", name, "k", k, flush=True)
        want = context_diagnostics(pkg, g, k)
        phi, c, mu = device_state(g)
        got = pkg.diagnostics_noalloc(g["pd"], float(g["sc0"][9]), phi.data_ptr(), c.data_ptr(), mu.data_ptr(),
                                      cells_per_thread=k)
        assert got == want, (name, got, want)


def test_noalloc_rejects_oversized_domain(pkg):
    g = load_golden("small_48x40")
    pd = dict(g["pd"])
    pd["nx"], pd["ny"] = 16384, 8192  # 64 x 8192 tiles at one row per tile
    phi, c, mu = device_state(g)  # never read: the size check comes first
    with pytest.raises(pkg.PearlError) as e:
        pkg.diagnostics_noalloc(pd, float(g["sc0"][9]), phi.data_ptr(), c.data_ptr(), mu.data_ptr(),
                                cells_per_thread=1)
    assert e.value.code == pkg.ERR_ARGUMENT and "exceeds the no-allocation scratch" in str(e.value)


def test_noalloc_reports_invalid_mixture(pkg):
    g = load_golden("small_48x40")
    phi0 = np.array(g["phi0"], dtype=np.float64)
    phi0.reshape(3, -1)[:, 17] = np.nan
    phi = torch.from_numpy(phi0).cuda()
    c = torch.from_numpy(np.ascontiguousarray(g["c0"], dtype=np.float64)).cuda()
    mu = torch.from_numpy(np.ascontiguousarray(g["mu0"], dtype=np.float64)).cuda()
    torch.cuda.synchronize()
    with pytest.raises(pkg.PearlError) as e:
        pkg.diagnostics_noalloc(g["pd"], float(g["sc0"][9]), phi.data_ptr(), c.data_ptr(), mu.data_ptr())
    assert e.value.code == pkg.ERR_MIXTURE and "Invalid mixture coefficients" in str(e.value)
    # a valid state right after: the flag of the failed call is not carried over
    g2 = load_golden("colony_256x128")
    want = context_diagnostics(pkg, g2, 0)
    p2, c2, m2 = device_state(g2)
    assert pkg.diagnostics_noalloc(g2["pd"], float(g2["sc0"][9]), p2.data_ptr(), c2.data_ptr(), m2.data_ptr()) == want
