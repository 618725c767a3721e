"""Tile-shape autotuner for batched ensembles.

A tile is 256 columns by k rows (k = cells per thread, one column per
thread).  The best k depends on the mix of domain sizes: small k gives more
tiles (more parallelism for short domains), larger k amortises the per-tile
halo rows and the block set-up over more cells.  The tuner measures instead
of modelling: for every candidate k it builds the batch, loads the given
states, runs a few warm-up steps per domain and then times a short run of
accepted steps, all on the device that will run the ensemble.  The
summation order of the blocked reductions depends on k, so the choice is
part of the run's reproducibility record (pearl_get_layout).
"""
from __future__ import annotations

import time

CANDIDATES = (1, 2, 4, 8, 16)


def tune_batch(pkg, params_list, mu_references, states, precision=0, candidates=CANDIDATES, steps=12, warmup=3,
               stop_time=1.0, device=0, log=None):
    """Return (best_k, {k: Mcell-steps/s}).

    states: one (phi, carbon, mu, scalars) per domain.  Each candidate starts
    from these states; the contexts are released before the next one is
    built, so the device holds one batch at a time."""
    cells = sum(int(p["nx"] if isinstance(p, dict) else p.nx) * int(p["ny"] if isinstance(p, dict) else p.ny)
                for p in params_list)
    n = len(params_list)
    rates = {}
    for k in candidates:
        b = pkg.Batch(params_list, mu_references, device=device, cells_per_thread=k, precision=precision)
        try:
            for d, (phi, c, mu, sc) in enumerate(states):
                b.set_domain_state(d, phi, c, mu, sc)
            r = b.run_all([stop_time] * n, [warmup] * n)
            if r["rc"] != 0:
                raise pkg.PearlError(r["rc"], f"warm-up failed for k={k}")
            b.counters()  # synchronises
            t0 = time.perf_counter()
            r = b.run_all([stop_time] * n, [steps] * n)
            wall = time.perf_counter() - t0
            if r["rc"] != 0:
                raise pkg.PearlError(r["rc"], f"timed run failed for k={k}")
            taken = sum(int(t) for t in r["taken"])
            rates[k] = cells * (taken / n) / wall / 1e6
        finally:
            b.close()
        if log:
            log(f"  autotune k={k}: {rates[k]:.1f} Mcell-steps/s")
    best = max(rates, key=rates.get)
    return best, rates
