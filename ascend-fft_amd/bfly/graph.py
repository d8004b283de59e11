"""hipGraph capture of library calls through the C ABI (bfly_graph_*).

    g = Graph()
    with torch.cuda.stream(s):          # a created stream, not the null one
        g.capture(lambda: lib.bfly_c2c(..., s.cuda_stream))
    g.replay(s)                         # re-runs the captured launches
    g.close()

Only device-pointer calls with their buffers allocated beforehand belong in
the captured function (no allocation, no synchronising helper).
"""
from __future__ import annotations

import ctypes as C

from . import _lib


class Graph:
    def __init__(self):
        self._exec = C.c_void_p()

    def capture(self, fn, stream: int) -> None:
        """Capture fn()'s launches on `stream` (the stream's handle)."""
        lib = _lib.hip()
        if self._exec.value:
            raise RuntimeError("graph already captured")
        _lib.check(lib.bfly_graph_begin(stream), "bfly_graph_begin")
        try:
            fn()
        finally:
            rc = lib.bfly_graph_end(stream, C.byref(self._exec))
        _lib.check(rc, "bfly_graph_end")

    def replay(self, stream: int) -> None:
        _lib.check(_lib.hip().bfly_graph_launch(self._exec, stream), "bfly_graph_launch")

    def close(self) -> None:
        if self._exec.value:
            _lib.hip().bfly_graph_destroy(self._exec)
            self._exec = C.c_void_p()
