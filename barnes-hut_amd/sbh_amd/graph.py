"""HIP graph capture of library calls (sbh_graph_* in include/sbh_amd.h).

    g = Graph()
    g.capture(lambda: gravity.gravity_force_batch(..., workspace=ws, out=f,
                                                  check_offsets=False))
    g.replay()          # re-runs every captured launch on the current stream

Captured calls must not allocate: pass preallocated workspaces and outputs.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional

import torch

from ._lib import call, stream_handle


class Graph:
    def __init__(self) -> None:
        self._exec = ctypes.c_void_p()

    @property
    def captured(self) -> bool:
        return bool(self._exec.value)

    def capture(self, fn: Callable[[], object], stream: Optional[torch.cuda.Stream] = None):
        """Records the launches `fn` makes on a side stream into the graph."""
        if self.captured:
            raise RuntimeError("graph already captured")
        side = stream or torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            h = stream_handle()
            call("sbh_graph_begin", h)
            try:
                result = fn()
            except BaseException:
                # end the capture and drop whatever was recorded; the error
                # of `fn` is the one the caller sees
                try:
                    call("sbh_graph_end", h, ctypes.byref(self._exec))
                except Exception:
                    pass
                self.close()
                raise
            call("sbh_graph_end", h, ctypes.byref(self._exec))
        torch.cuda.current_stream().wait_stream(side)
        return result

    def replay(self) -> None:
        if not self.captured:
            raise RuntimeError("nothing captured")
        call("sbh_graph_launch", self._exec, stream_handle())

    def close(self) -> None:
        if self.captured:
            call("sbh_graph_destroy", self._exec)
            self._exec = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
