"""The project's autotuner: it times candidate launch parameters of the library
(scatrs_set_tuning) on the caller's own workload and keeps the fastest, per device and workload
shape, in a JSON cache.

The parameters (keys of ``KEYS``):
  * ``sphere_waves``: persistent wavefronts per CU of the sphere batch;
  * ``node_waves``: persistent wavefronts per CU of the ensembles' sphere node solves;
  * ``chunk_share``: a chunk of ensembles stages its coefficients in 1/share of the free device
    memory (larger chunks wait less often for their slowest node, smaller ones leave memory);
  * ``split_blocks``: blocks per CU a small batch's angular products aim for;
  * ``sphere_plain`` (not tuned: a switch for tests): 0 runs every sphere of the sphere batch on
    the checked path instead of trying the branch-free plain path first;
  * ``sphere_wave`` (not tuned: set by a rule of the batch's size): when a sphere
    batch takes one wavefront per sphere (0 never, 1 always, f >= 2 when its terms times f are
    fewer than the GPU's lanes times its largest N).

Results do not depend on the first three; the fourth changes a small batch's summation order
(rounding only). Tuning coordinate by coordinate: each parameter in turn over its candidates,
the others held at their best so far, each candidate timed as the median of ``repeats`` calls
after one warm-up call.
"""

from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Callable, Sequence

KEYS = {"sphere_waves": 0, "node_waves": 1, "chunk_share": 2, "split_blocks": 3,
        "sphere_plain": 4, "sphere_wave": 5}
DEFAULTS = {"sphere_waves": 16, "node_waves": 8, "chunk_share": 4, "split_blocks": 2,
            "sphere_plain": 1, "sphere_wave": 8}
CANDIDATES = {
    "sphere_waves": (4, 6, 8, 12, 16, 20),
    "node_waves": (4, 6, 8, 12),
    "chunk_share": (2, 3, 4, 6),
    "split_blocks": (1, 2, 4, 8),
}

__all__ = ["CANDIDATES", "DEFAULTS", "KEYS", "TuningCache", "apply", "best_of", "tune"]


def best_of(timings: dict) -> object:
    """The candidate of smallest time (the first listed among equal times)."""
    return min(timings, key=lambda c: (timings[c], list(timings).index(c)))


def _median(v: list[float]) -> float:
    s = sorted(v)
    n = len(s)
    return s[n // 2] if n % 2 else 0.5 * (s[n // 2 - 1] + s[n // 2])


def tune(run: Callable[[], object], set_param: Callable[[str, int], None],
         params: Sequence[str], repeats: int = 3, sync: Callable[[], None] | None = None,
         candidates: dict | None = None, clock: Callable[[], float] = time.perf_counter) -> dict:
    """Tunes ``params`` (names of KEYS) for the workload ``run`` (one call of it, queued on the
    device; ``sync`` waits for the device). ``set_param(name, value)`` applies a value. Returns
    {"best": {name: value}, "timings": {name: {value: seconds}}}; the best values stay
    applied."""
    cand = dict(CANDIDATES, **(candidates or {}))
    sync = sync or (lambda: None)
    best = {p: DEFAULTS[p] for p in params}
    timings: dict = {}
    for p in params:
        row = {}
        for v in cand[p]:
            set_param(p, v)
            run()
            sync()
            samples = []
            for _ in range(repeats):
                t0 = clock()
                run()
                sync()
                samples.append(clock() - t0)
            row[v] = _median(samples)
        best[p] = best_of(row)
        set_param(p, best[p])
        timings[p] = row
    return {"best": best, "timings": timings}


def apply(ctx, values: dict) -> None:
    """Sets tuned values on a library context (scatrs_amd._native.Context)."""
    for name, v in values.items():
        ctx.set_tuning(KEYS[name], int(v))


class TuningCache:
    """Tuned values per (device, workload key), in a JSON file (default
    $SCATRS_TUNING_CACHE or ~/.cache/scatrs_amd/tuning.json)."""

    def __init__(self, path: str | os.PathLike | None = None) -> None:
        default = Path(os.environ.get("SCATRS_TUNING_CACHE",
                                      Path.home() / ".cache" / "scatrs_amd" / "tuning.json"))
        self.path = Path(path) if path is not None else default
        try:
            self.data = json.loads(self.path.read_text())
        except (OSError, ValueError):
            self.data = {}

    @staticmethod
    def key(device: str, workload: str) -> str:
        return f"{device}|{workload}"

    def get(self, device: str, workload: str) -> dict | None:
        return self.data.get(self.key(device, workload))

    def put(self, device: str, workload: str, values: dict) -> None:
        self.data[self.key(device, workload)] = dict(values)
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            tmp = self.path.with_suffix(".tmp")
            tmp.write_text(json.dumps(self.data, indent=1, sort_keys=True))
            os.replace(tmp, self.path)
        except OSError:
            pass
