"""Launch-shape autotuner for the library's tunable kernels.

The library exposes, per tunable operation and host thread, a variant index
(``ap_tune_count/get/set``): the per-ray forward projector's pixels per
workgroup, the column forward projector's detector rows per workgroup, and the
voxel-driven backprojector's voxels per workgroup.  All variants of one
operation compute bit-identical results; only the decomposition differs, and
which is fastest depends on the problem (detector and volume size, view count).

``Autotuner.tune(op, key, run)`` times every variant of ``op`` on the caller's
``run()`` (which must be safe to repeat: it writes scratch buffers only) with
HIP events on the current stream, and records the fastest for the problem
class ``key``.  ``Autotuner.select(op, key)`` later sets the recorded variant
before a launch -- also while a HIP graph is captured, since the variant is read
when the launch is recorded.  Results can be saved to and loaded from JSON.
"""

from __future__ import annotations

import json
import pathlib
from collections.abc import Callable, Hashable

from . import _abi

OPS = {
    "project_ray": _abi.TUNE_PROJECT_RAY,
    "project_column": _abi.TUNE_PROJECT_COLUMN,
    "backproject": _abi.TUNE_BACKPROJECT,
}


def variant_count(op: int) -> int:
    n = _abi.lib().ap_tune_count(op)
    if n < 0:
        msg = f"unknown tuning op {op}"
        raise ValueError(msg)
    return n


def get_variant(op: int) -> int:
    return int(_abi.lib().ap_tune_get(op))


def set_variant(op: int, variant: int) -> None:
    _abi.check(_abi.lib().ap_tune_set(op, variant), "tune_set")


def size_class(*sizes: int) -> tuple[int, ...]:
    """Bucket sizes to powers of two (rounded up): problems of one class share
    a tuning result."""
    return tuple(1 << max(0, int(s) - 1).bit_length() for s in sizes)


class Autotuner:
    """Times launch-shape variants per (operation, problem class)."""

    def __init__(self, *, repeats: int = 3) -> None:
        self.repeats = repeats
        self._best: dict[tuple[int, Hashable], int] = {}
        self._times: dict[tuple[int, Hashable], list[float]] = {}

    def known(self, op: int, key: Hashable) -> bool:
        return (op, key) in self._best

    def best(self, op: int, key: Hashable) -> int | None:
        return self._best.get((op, key))

    def times_ms(self, op: int, key: Hashable) -> list[float] | None:
        return self._times.get((op, key))

    def tune(self, op: int, key: Hashable, run: Callable[[], None]) -> int:
        """Time every variant of ``op`` on ``run`` (once unless ``key`` is new)
        and leave the fastest selected; returns it."""
        import torch  # noqa: PLC0415

        if (op, key) in self._best:
            return self.select(op, key)
        stream = torch.cuda.current_stream()
        times = []
        for var in range(variant_count(op)):
            set_variant(op, var)
            run()   # warm-up (code objects, caches)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(self.repeats):
                run()
            e1.record(stream)
            e1.synchronize()
            times.append(e0.elapsed_time(e1) / self.repeats)
        best = min(range(len(times)), key=times.__getitem__)
        self._best[(op, key)] = best
        self._times[(op, key)] = times
        set_variant(op, best)
        return best

    def select(self, op: int, key: Hashable) -> int:
        """Set the recorded variant for ``key`` (variant 0 if none is recorded)."""
        var = self._best.get((op, key), 0)
        set_variant(op, var)
        return var

    # ---- persistence ------------------------------------------------------
    def save(self, path) -> None:
        data = [{"op": op, "key": list(key) if isinstance(key, tuple) else key, "variant": var,
                 "times_ms": self._times.get((op, key))} for (op, key), var in self._best.items()]
        pathlib.Path(path).write_text(json.dumps(data, indent=1))

    def load(self, path) -> None:
        for e in json.loads(pathlib.Path(path).read_text()):
            key = tuple(e["key"]) if isinstance(e["key"], list) else e["key"]
            if not 0 <= e["variant"] < variant_count(e["op"]):
                msg = f"stored variant {e['variant']} out of range for op {e['op']}"
                raise ValueError(msg)
            self._best[(e["op"], key)] = e["variant"]
            if e.get("times_ms") is not None:
                self._times[(e["op"], key)] = e["times_ms"]
