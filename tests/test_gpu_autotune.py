"""Tile-height autotuner for batched ensembles (autotune.tune_batch)."""
import importlib
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_candidate_and_batches_reproduce(pkg):
    autotune = importlib.import_module("pearl_amd.autotune")
    gs = [load_golden("small_48x40"), load_golden("colony_256x128")]
    params = [g["pd"] for g in gs]
    mus = [float(g["sc0"][9]) for g in gs]
    states = [(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9]) for g in gs]
    print("autotune", flush=True)
    sys.stdout.flush()
    best, rates = autotune.tune_batch(pkg, params, mus, states, candidates=(1, 4, 16), steps=4, warmup=2)
    assert best in (1, 4, 16) and set(rates) == {1, 4, 16}
    assert all(v > 0 for v in rates.values()) and rates[best] == max(rates.values())
    # a batch at the chosen tile height steps each domain exactly like a
    # single context with that tile height
    with pkg.Batch(params, mus, cells_per_thread=best) as b:
        for d, st in enumerate(states):
            b.set_domain_state(d, *st)
        r = b.run_all([1.0, 1.0], [6, 6])
        assert r["rc"] == 0
        got = [b.get_domain_state(d) for d in range(2)]
    for d, g in enumerate(gs):
        with pkg.Solver(g["pd"], mus[d], cells_per_thread=best) as s:
            s.set_state(*states[d])
            s.run(1.0, 6)
            ref = s.get_state()
        for x, y in zip(got[d][:3], ref[:3]):
            assert np.array_equal(x, y)
