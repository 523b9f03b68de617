"""Launch timing of one kernel inside a timed region.

``KernelClock(name)`` arms the library's launch clock (``smx_clock_start``):
HIP events are recorded on the launch stream immediately before and after
every launch of that entry point, from the C side, and the kernel itself
counts the algorithmic units it processes in a device counter.  Nothing is
added on the host per launch and nothing synchronises inside the timed
region; ``stop`` collects the spans and the count.
"""
import ctypes

import torch

from . import runtime

TAGS = {"fg_post": 1, "pc_apply": 2}
_armed = [False]


def armed():
    """True while a clock is running: launches then record events from the C
    side, which must not be captured into HIP graphs."""
    return _armed[0]


class KernelClock:
    def __init__(self, name, capacity=60000):
        if name not in TAGS:
            raise ValueError("no launch clock for %r (have %s)" % (name, sorted(TAGS)))
        self.name, self.capacity = name, int(capacity)
        self.launches, self.total_ms, self.total_units = 0, 0.0, 0

    def start(self):
        runtime.load().clock_start(TAGS[self.name], self.capacity)
        _armed[0] = True

    def stop(self):
        torch.cuda.synchronize()
        ms, n, u = ctypes.c_double(0.0), ctypes.c_int(0), ctypes.c_longlong(0)
        runtime.load().clock_stop(ctypes.byref(ms), ctypes.byref(n), ctypes.byref(u))
        _armed[0] = False
        if n.value >= self.capacity:
            raise RuntimeError("launch clock full: raise its capacity above %d" % self.capacity)
        self.launches += n.value
        self.total_ms += ms.value
        self.total_units += u.value

    def summary(self):
        """-> (launches, mean launch ms, total units) over every start/stop."""
        if not self.launches:
            return 0, 0.0, 0
        return self.launches, self.total_ms / self.launches, self.total_units
