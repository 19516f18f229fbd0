"""HIP-graph capture of launch sequences (MAP / leapfrog steps).

Every entry point of the C ABI launches on the stream it is given and neither
allocates nor synchronises, so a sequence of calls made on torch's current
stream can be recorded into one HIP graph (``torch.cuda.CUDAGraph`` is a HIP
graph on ROCm) and replayed.  Scalars passed by value (step counts, learning
rates) are baked into the recorded launches; buffers are reused in place.
No call inside a capture may clear memory with hipMemsetAsync -- none of the
recorded entry points does.
"""
from __future__ import annotations

from typing import Callable

import torch


def capture(fn: Callable[[], None]) -> torch.cuda.CUDAGraph:
    """Record the launches ``fn`` issues on the current stream into a graph
    (not executed during the capture).  ``graph.replay()`` runs them."""
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        fn()
    return g
