"""Launch timing of kernels inside a timed region.

``KernelClock(names)`` arms the library's launch clock (``smx_clock_start``)
for one or more entry points: HIP events are recorded on the launch stream
immediately before and after every launch of each, from the C side, and the
kernels themselves count the algorithmic units they process in a device
counter per entry point.  Nothing is added on the host per launch and
nothing synchronises inside the timed region; ``stop`` collects the spans
and the counts.
"""
import ctypes

import torch

from . import runtime

TAGS = {"fg_post": 1, "pc_apply": 2, "dense_solve": 3}
# entry points launched inside the FGMRES rounds that HIP graphs capture; the
# dense solves run only in eager rescue solves
_IN_GRAPH_ROUNDS = ("fg_post", "pc_apply")
_armed = [False]


def armed():
    """True while a clock runs on an entry point of the graph-captured FGMRES
    rounds: its launches then record events from the C side, which must not
    be captured into HIP graphs, so those rounds run eagerly."""
    return _armed[0]


class KernelClock:
    def __init__(self, names, capacity=60000):
        names = (names,) if isinstance(names, str) else tuple(names)
        for n in names:
            if n not in TAGS:
                raise ValueError("no launch clock for %r (have %s)" % (n, sorted(TAGS)))
        self.names, self.capacity = names, int(capacity)
        self.stats = {n: [0, 0.0, 0] for n in names}   # launches, total ms, units

    def start(self):
        mask = 0
        for n in self.names:
            mask |= 1 << TAGS[n]
        runtime.load().clock_start(mask, self.capacity)
        _armed[0] = any(n in _IN_GRAPH_ROUNDS for n in self.names)

    def stop(self):
        torch.cuda.synchronize()
        lib = runtime.load()
        try:
            for n in self.names:
                ms, c, u = ctypes.c_double(0.0), ctypes.c_int(0), ctypes.c_longlong(0)
                lib.clock_stop(TAGS[n], ctypes.byref(ms), ctypes.byref(c), ctypes.byref(u))
                if c.value >= self.capacity:
                    raise RuntimeError("launch clock of %s full: raise its capacity above %d"
                                       % (n, self.capacity))
                st = self.stats[n]
                st[0] += c.value
                st[1] += ms.value
                st[2] += u.value
        finally:
            _armed[0] = False

    def summary(self, name=None):
        """-> (launches, mean launch ms, total units) of ``name`` (default: the
        first clocked entry point) over every start/stop."""
        n, ms, u = self.stats[self.names[0] if name is None else name]
        if not n:
            return 0, 0.0, 0
        return n, ms / n, u
