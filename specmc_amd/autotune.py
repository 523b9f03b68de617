"""Autotuner: measured kernel choices for the shapes an engine runs.

``tune_gemms(eng, step)`` records every GEMM that one call of ``step()``
issues on ``eng`` (shape, batch, operand layout), times each kernel choice of
the library (vector-FMA tiles 128x96 ... 32x32, matrix-core 64x64 and 32x32
with chunked or single-chunk K staging) on scratch operands of the same
shape and strides with HIP events, and installs the fastest in the library's
table (``smx_gemm_tune``).  Choices are keyed by (precision, M, N, K, batch,
operand layout), process-wide; they change products only at roundoff.

``tune_pc(eng, iters)`` chooses how the per-mode preconditioner blocks are
applied -- an LU with two triangular sweeps per apply, or an explicit inverse
applied as one matrix-vector product -- from the measured factor and apply
times and the expected number of applies per factorisation.

``tune(eng, step)`` runs both and returns a report.
"""
import numpy as np
import torch

CODES = (10, 11, 12, 0, 1, 2, 3, 4, 5, 6)
CODE_NAMES = {0: "fma128x96", 1: "fma128x64", 2: "fma64x96", 3: "fma64x64", 4: "fma64x32",
              5: "fma32x64", 6: "fma32x32", 10: "mfma64x64", 11: "mfma32x32",
              12: "mfma32x32-1chunk"}


def _extent(rows, cols, rs, cs, n1, s1, n2, s2):
    return 1 + (rows - 1) * rs + (cols - 1) * cs + (n1 - 1) * s1 + (n2 - 1) * s2


class _Recorder:
    def __init__(self, eng):
        self.eng, self.calls = eng, {}
        self.orig = eng.gemm

    def __enter__(self):
        def rec(M, N, K, A, ars, acs, Bm, brs, bcs, C, crs, ccs, **kw):
            nz1, nz2 = kw.get("nz1", 1), kw.get("nz2", 1)
            key = (self.eng.prec, M, N, K, nz1 * nz2, int(acs == 1), int(bcs == 1))
            if key not in self.calls:
                sa, sb, sc = kw.get("sa", (0, 0)), kw.get("sb", (0, 0)), kw.get("sc", (0, 0))
                self.calls[key] = dict(M=M, N=N, K=K, ars=ars, acs=acs, brs=brs, bcs=bcs,
                                       crs=crs, ccs=ccs, nz1=nz1, nz2=nz2, sa=sa, sb=sb,
                                       sc=sc, beta=kw.get("beta", 0.0), role=kw.get("role", 0),
                                       count=0)
            self.calls[key]["count"] += 1
            return self.orig(M, N, K, A, ars, acs, Bm, brs, bcs, C, crs, ccs, **kw)
        self.eng.gemm = rec
        return self

    def __exit__(self, *exc):
        del self.eng.gemm          # back to the class method
        return False


def _time_call(eng, c, code, reps):
    L = eng.L
    dt = eng.dtype
    dev = eng.dev
    ea = _extent(c["M"], c["K"], c["ars"], c["acs"], c["nz1"], c["sa"][0], c["nz2"], c["sa"][1])
    eb = _extent(c["K"], c["N"], c["brs"], c["bcs"], c["nz1"], c["sb"][0], c["nz2"], c["sb"][1])
    ec = _extent(c["M"], c["N"], c["crs"], c["ccs"], c["nz1"], c["sc"][0], c["nz2"], c["sc"][1])
    A = torch.rand(ea, dtype=dt, device=dev)
    B = torch.rand(eb, dtype=dt, device=dev)
    C = torch.zeros(ec, dtype=dt, device=dev)
    key = (eng.prec, c["M"], c["N"], c["K"], c["nz1"] * c["nz2"], int(c["acs"] == 1),
           int(c["bcs"] == 1))
    L.gemm_tune(*key, code)
    s = eng.stream()

    def go():
        L.gemm_ex(c["role"], eng.prec, c["M"], c["N"], c["K"], 1.0, c["beta"],
                  A.data_ptr(), c["ars"], c["acs"], c["sa"][0], c["sa"][1],
                  B.data_ptr(), c["brs"], c["bcs"], c["sb"][0], c["sb"][1],
                  C.data_ptr(), c["crs"], c["ccs"], c["sc"][0], c["sc"][1],
                  c["nz1"], c["nz2"], None, 1, None, 0, 1, s)
    go()
    go()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        go()
    e1.record()
    e1.synchronize()
    return 1e3 * e0.elapsed_time(e1) / reps          # us per launch


def tune_gemms(eng, step, reps=8, codes=CODES):
    """-> {shape key: dict(code, us, default_us, calls_per_step)}."""
    with _Recorder(eng) as rec:
        step()
    torch.cuda.synchronize()
    L = eng.L
    out = {}
    for key, c in rec.calls.items():
        L.gemm_tune(*key, -1)
        times = {}
        for code in codes:
            if code == 12 and c["K"] > 160:
                continue
            times[code] = _time_call(eng, c, code, reps)
        best = min(times, key=times.get)
        L.gemm_tune(*key, best)
        out[key] = dict(code=best, kernel=CODE_NAMES[best], us=times[best],
                        candidates={CODE_NAMES[k]: v for k, v in times.items()},
                        calls_per_step=c["count"], role=c["role"])
    torch.cuda.synchronize()
    return out


def tune_pc(eng, iters):
    """Pick 'lu' or 'inv' for the per-mode preconditioner blocks of ``eng``
    from measured times: factor once + ``iters`` applies."""
    B, Nk, N = eng.B, eng.Nk, eng.Nz
    dev, dt = eng.dev, eng.dtype
    rng = np.random.default_rng(0)
    blk = rng.normal(size=(N, N)) + N * np.eye(N)
    mats = torch.tensor(np.ascontiguousarray(blk.T), dtype=dt, device=dev).reshape(1, 1, N * N)
    mats = mats.expand(B, Nk, N * N).contiguous()
    piv = torch.zeros(B, Nk, N, dtype=torch.int32, device=dev)
    vin = torch.rand(B, Nk * N, dtype=dt, device=dev)
    zout = torch.zeros_like(vin)
    ev = lambda: torch.cuda.Event(enable_timing=True)
    res = {}
    from .engine import inverse_fits_lds
    keep = eng.pc_apply
    modes = ("inv", "lu") if inverse_fits_lds(N, eng.esize) else ("lu",)
    for mode in modes:
        eng.pc_apply = mode
        work = mats.clone()
        eng._pc_factor(work, piv, None)
        e0, e1, e2 = ev(), ev(), ev()
        work = mats.clone()
        e0.record()
        eng._pc_factor(work, piv, None)
        e1.record()
        for _ in range(5):
            eng._pc_apply(work, piv, vin, zout, 1.0, None)
        e2.record()
        e2.synchronize()
        f, a = e0.elapsed_time(e1), e1.elapsed_time(e2) / 5
        res[mode] = dict(factor_ms=f, apply_ms=a, total_ms=f + iters * a)
    eng.pc_apply = keep
    best = min(res, key=lambda m: res[m]["total_ms"])
    eng.set_pc_apply(best)
    return dict(choice=best, iters=iters, **res)


def tune(eng, step, iters=None, reps=8):
    """GEMM kernels for the shapes one ``step()`` issues, then the
    preconditioner-block schedule (``iters`` applies per factorisation,
    default: the Stokes iterations of the last solve, at least 10)."""
    g = tune_gemms(eng, step, reps=reps)
    if iters is None:
        it = getattr(eng, "last_bfbt_iters", None)
        iters = max(10, int(np.max(it)) if it is not None and len(it) else 10)
    p = tune_pc(eng, iters) if eng.kind != "const" else None
    return dict(gemm={"%s" % (k,): v for k, v in g.items()}, pc=p)
