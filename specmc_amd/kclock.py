"""HIP-event timing of selected kernel launches on the engine's stream.

An ``Engine`` with ``clock`` set records an event pair around every launch of
the named kernels, together with the number of units (systems, rows) the
launch processes as a device-side count, so a benchmark can report a
kernel's average launch duration and its achieved bandwidth from the same
timed region without host synchronisation inside it.
"""
import torch


class KernelClock:
    def __init__(self, names):
        self.names = set(names)
        self.spans = {k: [] for k in self.names}
        self.units = {k: [] for k in self.names}

    def wants(self, name):
        return name in self.names

    def begin(self):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return ev

    def end(self, name, ev0, units):
        """``units``: int, or a device tensor whose sum is the unit count."""
        ev1 = torch.cuda.Event(enable_timing=True)
        ev1.record()
        self.spans[name].append((ev0, ev1))
        self.units[name].append(units if isinstance(units, int) else units.sum())

    def summary(self, name):
        """-> (launches, mean launch ms, total units) after a synchronize."""
        torch.cuda.synchronize()
        sp = self.spans[name]
        if not sp:
            return 0, 0.0, 0
        ms = sum(a.elapsed_time(b) for a, b in sp)
        u = 0
        dev = [x for x in self.units[name] if not isinstance(x, int)]
        if dev:
            u += int(torch.stack([d.to(torch.int64) for d in dev]).sum().item())
        u += sum(x for x in self.units[name] if isinstance(x, int))
        return len(sp), ms / len(sp), u
