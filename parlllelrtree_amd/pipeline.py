"""The map-overlay pipeline on device tensors: polygon MBRs -> R-tree over
layer B's MBRs -> join with layer A's MBRs -> batched FP64 clip of every
candidate pair -> total area and non-empty count, run eagerly or captured
once into a hipGraph and replayed (prt_pipeline_* of libprt).

The reference's stages are separate programs (BotomUp-rtree.cu main BU:642,
TopDown-rtree.cu main TD:483) feeding run() (vector_map_overlay_gpc_kernel.cu
:61-79), which counts non-empty results as count1 (OV:45-58).
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import PipelineIO, PrtError, TuneStruct, check, lib
from .rtree import PolyDev

PRT_E_OVERFLOW = -4
TREES = {"bu": 0, "td": 1}


def _stream_ptr(s: torch.cuda.Stream | None = None) -> ctypes.c_void_p:
    s = s if s is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class OverlayPipeline:
    """Buffers and (optionally) a captured graph of one overlay pipeline.

    `cap` bounds the pair list; `result()` raises PrtError(PRT_E_OVERFLOW)
    when the join found more.  The graph bakes in the addresses of A, B and
    of this object's buffers: a replay reads whatever they then hold."""

    def __init__(self, A: PolyDev, B: PolyDev, cap: int, fanout: int = 16, tree: str = "bu",
                 tune: TuneStruct | None = None):
        if A.xy.dtype != torch.float64 or B.xy.dtype != torch.float64:
            raise ValueError("the pipeline clips FP64 polygons")
        dev = A.xy.device
        self.A, self.B, self.cap, self.tune = A, B, int(cap), tune
        self.pairs = torch.empty((max(self.cap, 1), 2), dtype=torch.int32, device=dev)
        self.area = torch.empty(max(self.cap, 1), dtype=torch.float64, device=dev)
        self.count = torch.zeros(2, dtype=torch.int64, device=dev)
        self.total = torch.zeros(1, dtype=torch.float64, device=dev)
        self.nonempty = torch.zeros(1, dtype=torch.int64, device=dev)
        io = PipelineIO()
        io.xyA, io.offA, io.nA = A.xy.data_ptr(), A.off.data_ptr(), A.off.numel() - 1
        io.xyB, io.offB, io.nB = B.xy.data_ptr(), B.off.data_ptr(), B.off.numel() - 1
        io.max_verts = A.max_verts + B.max_verts
        io.fanout, io.tree, io.cap = fanout, TREES[tree], self.cap
        io.pairs, io.area, io.count = (self.pairs.data_ptr(), self.area.data_ptr(),
                                       self.count.data_ptr())
        io.total, io.nonempty = self.total.data_ptr(), self.nonempty.data_ptr()
        self.io = io
        need = lib().prt_pipeline_workspace_size(ctypes.byref(io))
        if need == 0:
            raise PrtError(-1, lib().prt_last_error().decode())
        self.ws = torch.empty(need, dtype=torch.uint8, device=dev)
        self._graph = ctypes.c_void_p(0)

    def _tp(self):
        return ctypes.byref(self.tune) if self.tune is not None else None

    def run(self) -> None:
        """Enqueue the pipeline on the current stream (eager launches)."""
        check(lib().prt_pipeline_run(ctypes.byref(self.io), self._tp(), self.ws.data_ptr(),
                                     self.ws.numel(), _stream_ptr()))

    def capture(self) -> None:
        """Record the pipeline into a hipGraph (nothing runs yet)."""
        self.close()
        side = torch.cuda.Stream(self.ws.device)
        h = ctypes.c_void_p(0)
        check(lib().prt_pipeline_capture(ctypes.byref(self.io), self._tp(), self.ws.data_ptr(),
                                         self.ws.numel(), _stream_ptr(side), ctypes.byref(h)))
        self._graph = h

    @property
    def captured(self) -> bool:
        return bool(self._graph.value)

    def replay(self) -> None:
        """Launch the captured graph on the current stream."""
        if not self.captured:
            raise RuntimeError("capture() first")
        check(lib().prt_pipeline_launch(self._graph, _stream_ptr()))

    def result(self):
        """(pairs (n, 2) int32 view as (A id, B id), area (n,), total, nonempty)
        of the last run or replay (synchronises)."""
        found, status = (int(v) for v in self.count.cpu())
        if found > self.cap:
            raise PrtError(PRT_E_OVERFLOW, f"pipeline: {found} pairs > capacity {self.cap}")
        if status:
            raise PrtError(PRT_E_OVERFLOW, "pipeline: a query had more hits than the device "
                                           "segment sort takes")
        return (self.pairs[:found], self.area[:found], float(self.total.item()),
                int(self.nonempty.item()))

    def close(self) -> None:
        if self.captured:
            check(lib().prt_pipeline_destroy(self._graph))
            self._graph = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
