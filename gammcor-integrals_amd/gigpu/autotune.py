"""Autotuner of the ERI launch layer and the batched form.

The ERI launches choose their quartet tiles by rule (ket pairs stacked per
MFMA tile, a floor on the number of tiles, which classes go to the
lane-per-quartet kernels) and the full-dimension vectors split their
triangular GEMM into block rows; all of these are options of a context
(`gi_set_option`).  `tune` times the real workload -- one decomposition, or
the batched decomposition of many systems -- over candidate values of each
option in turn (coordinate descent from the current defaults, the best value
kept before the next option is tried) and returns the best settings, which
`apply` sets on any number of contexts.  The results of a decomposition do
not depend on these settings beyond floating-point rounding; the pivot count
is checked to stay the same on every trial.
"""
from __future__ import annotations

import time

from . import lib

# candidate values per option, defaults first
CANDIDATES = {
    "eri_max_kt": [32, 16, 64],
    "eri_min_tiles": [8192, 4096, 16384],
    "eri_lane_emax": [4, 2],
    "vec_blocks": [20, 12, 30],
}
DEFAULTS = {k: v[0] for k, v in CANDIDATES.items()}


def apply(contexts, settings: dict) -> None:
    for c in contexts:
        for k, v in settings.items():
            if k != "nworkers":
                c.set_option(k, v)


def _timed(run, reps):
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        out = run()
        best = min(best, time.perf_counter() - t0)
    return best, out


def tune(contexts, tau: float = 1e-6, candidates: dict | None = None, nworkers: list | None = None,
         reps: int = 2, log=None):
    """Coordinate-descent tuning on `contexts` (each with its basis set).  One
    context tunes gi_chol_build; several tune gi_chol_build_batch, and then
    `nworkers` (a list of worker counts) is tuned too.  Returns
    (settings, seconds, trials)."""
    candidates = dict(CANDIDATES if candidates is None else candidates)
    batched = len(contexts) > 1
    best = {k: v[0] for k, v in candidates.items()}
    if batched:
        best["nworkers"] = (nworkers or [8])[0]

    def run(settings):
        apply(contexts, settings)
        if batched:
            lib.chol_build_batch(contexts, tau=tau, nworkers=settings["nworkers"])
        else:
            contexts[0].chol_build(tau=tau)
        return [c.chol_dims()["nvecs"] for c in contexts]

    run(best)                                    # warm-up: code objects, pair tables, pools
    t_best, ref = _timed(lambda: run(best), reps)
    trials = [(dict(best), t_best)]
    axes = list(candidates.items())
    if batched and nworkers:
        axes.append(("nworkers", list(nworkers)))
    for name, values in axes:
        for v in values:
            if v == best[name]:
                continue
            trial = dict(best)
            trial[name] = v
            t, nv = _timed(lambda: run(trial), reps)
            if nv != ref:
                raise RuntimeError("autotune: %s=%s changed the decomposition" % (name, v))
            trials.append((dict(trial), t))
            if log:
                log("%s=%s: %.4f s (best %.4f s)" % (name, v, t, t_best))
            if t < t_best:
                best, t_best = trial, t
    apply(contexts, best)
    return best, t_best, trials
