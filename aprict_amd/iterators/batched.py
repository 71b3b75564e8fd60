"""Many independent SIRT reconstructions in one HIP graph launch.

``BatchedSIRT`` holds one ``SIRT_flatpanel`` per problem (each with its own
volume, measurements, preconditioner and residual buffer; sizes may all
differ).  ``capture(iterations)`` records ``iterations`` SIRT updates of every
problem into one HIP graph: the problems are spread over several streams by
longest-processing-time first on ``N^3 V``, each stream runs its problems one
after the other, and the streams fork from and join the capture stream, so the
graph lets independent problems overlap on the GPU.  Each SIRT update is a
single library call (``ap_sirt_iterate``: residual fused into the forward
projector, preconditioner multiply in the backprojector), with no host
synchronisation, allocation or memset, so it can be captured.  ``run()``
replays the graph: one launch performs all the updates.

With an ``Autotuner``, each problem's launch shapes (forward projector and
backprojector) are timed per problem class before capture, and the chosen
variants are set before each problem's calls are recorded.

This batched driver has no counterpart in the reference, which solves one
problem per ``SIRT_flatpanel`` (``iterators/sirt_vram.py:62-149``).
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

from .. import _abi
from ..autotune import Autotuner, size_class
from ..cb import flatpanel as cbfp
from ..volume import Volume
from .sirt_vram import SIRT_flatpanel


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _is_upright(pg: cbfp.ProjGeom) -> np.ndarray:
    v = np.broadcast_to(pg.det_vvec, pg.projs_list_shape + (3,)).reshape(-1, 3)
    return (v[:, 0] == 0) & (v[:, 1] == 0) & (v[:, 2] != 0)


class BatchedSIRT:
    """SIRT on a list of independent (volume, measurements) problems."""

    def __init__(self, problems: Sequence[tuple[Volume, cbfp.ProjectionSet]],
                 config: SIRT_flatpanel.Config | None = None, *, streams: int = 4,
                 tuner: Autotuner | None = None) -> None:
        if not problems:
            msg = "BatchedSIRT needs at least one problem"
            raise ValueError(msg)
        config = config or SIRT_flatpanel.Config()
        self.solvers = [SIRT_flatpanel(config, v, p) for v, p in problems]
        self.nstreams = max(1, int(streams))
        self.tuner = tuner
        self._graph = None
        self._graph_iters = 0
        self._keys = [self._problem_keys(s) for s in self.solvers]
        if tuner is not None:
            for s, keys in zip(self.solvers, self._keys, strict=True):
                self._tune(s, keys)

    # ---- tuning -------------------------------------------------------------
    @staticmethod
    def _problem_keys(s: SIRT_flatpanel) -> dict[int, tuple]:
        vol, meas = s.volume, s._meas
        pg = meas.proj_geom
        nu, nv = int(np.max(pg.nu)), int(np.max(pg.nv))
        views = int(np.prod(pg.projs_list_shape))
        cls = size_class(max(vol.shape), max(nu, nv), views) + (str(vol.data.dtype),)
        upright = _is_upright(pg)
        keys = {_abi.TUNE_BACKPROJECT: cls}
        if upright.any():
            keys[_abi.TUNE_PROJECT_COLUMN] = cls
        if not upright.all():
            keys[_abi.TUNE_PROJECT_RAY] = cls
        return keys

    def _tune(self, s: SIRT_flatpanel, keys: dict[int, tuple]) -> None:
        torch = _torch()
        scratch_p = s._proj_scratch_space   # overwritten by the next iterate()
        scratch_v = s.volume.zeroed_copy()
        cfg_fp = cbfp.ConfigProject()
        cfg_bp = s._config_bp
        for op, key in keys.items():
            if op == _abi.TUNE_BACKPROJECT:
                def run(s=s, cfg_bp=cfg_bp, scratch_v=scratch_v, scratch_p=scratch_p):
                    cbfp.backproject(cfg_bp, scratch_v, scratch_p, summand_multiplier=s.preconditioner)
            else:
                def run(s=s, cfg_fp=cfg_fp, scratch_p=scratch_p):
                    cbfp.project(cfg_fp, s.volume, scratch_p)
            self.tuner.tune(op, key, run)
        torch.cuda.synchronize()
        del scratch_v

    def _select(self, i: int) -> None:
        if self.tuner is not None:
            for op, key in self._keys[i].items():
                self.tuner.select(op, key)

    # ---- execution ----------------------------------------------------------
    def _assignment(self) -> list[list[int]]:
        """Problems per stream, longest first onto the least-loaded stream."""
        cost = []
        for s in self.solvers:
            pg = s._meas.proj_geom
            views = int(np.prod(pg.projs_list_shape))
            cost.append(float(np.prod(s.volume.shape)) * views)
        lanes = [[] for _ in range(self.nstreams)]
        load = [0.0] * self.nstreams
        for i in sorted(range(len(cost)), key=lambda i: -cost[i]):
            k = min(range(self.nstreams), key=load.__getitem__)
            lanes[k].append(i)
            load[k] += cost[i]
        return [lane for lane in lanes if lane]

    def run_eager(self, iterations: int) -> None:
        """``iterations`` updates of every problem, launched one by one on the
        current stream."""
        for i, s in enumerate(self.solvers):
            self._select(i)
            for _ in range(iterations):
                s.iterate()

    def capture(self, iterations: int) -> None:
        """Record ``iterations`` updates of every problem into one HIP graph."""
        torch = _torch()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        lanes = self._assignment()
        streams = [torch.cuda.Stream() for _ in lanes]
        with torch.cuda.graph(graph):
            origin = torch.cuda.current_stream()
            for st in streams:
                st.wait_stream(origin)
            for st, lane in zip(streams, lanes, strict=True):
                with torch.cuda.stream(st):
                    for i in lane:
                        self._select(i)
                        for _ in range(iterations):
                            self.solvers[i].iterate()
            for st in streams:
                origin.wait_stream(st)
        self._graph = graph
        self._graph_iters = iterations

    @property
    def captured_iterations(self) -> int:
        return self._graph_iters

    def run(self) -> None:
        """Replay the captured graph: ``captured_iterations`` more updates of
        every problem in one launch."""
        if self._graph is None:
            msg = "call capture(iterations) first"
            raise RuntimeError(msg)
        self._graph.replay()

    @property
    def volumes(self) -> list[Volume]:
        return [s.volume for s in self.solvers]
