"""Request sets of the benchmark workloads (seeded, reproducible).

* ``bench_64x96_config``       -- 64x96 eta_T, deta_T = 1e3, Ra = 1e4, unit box (Lx = 2
  on the mirror-doubled carrier), BFBT route; theta0 = conductive +
  3e-2 cos(pi x) sin(pi z); tstop = 5e-3.
* ``heated_2to1_requests``     -- the heated time-dependent case: Lx = 4 (2:1 box),
  64x96, Ra = 1.25e4, H = 1.5, eta_T deta_T = 1e3, theta0 = heated conductive
  profile + seeded low modes (``default_rng(2)``); march to t = 0.03 on the
  legacy (default) or BFBT Stokes route.
* ``ensemble_requests`` -- 256 runs from ``default_rng(3)``: grids uniform
  over {32x24, 64x48, 64x96, 128x96}, Ra = 10^U(4,6), deta_T = 10^U(0,4),
  model uniform over {eta_z, eta_T, arrhenius (Ts = 1, E = 2 ln deta_T)},
  H ~ U(0,2), Di = 0, Lx uniform over {2, 4}, seeded theta0 with its own H,
  all arriving at t = 0.
* ``stream_requests`` -- Poisson arrivals from ``default_rng(4)``: grids
  uniform over {32x24, 64x48, 64x96, 128x96, 128x192, 256x192}, arrhenius
  (Ts = 1) with deta_T = 10^U(2,5) and the window [1e-5, 1], Ra = 10^U(4,7),
  H ~ U(0,2), Di ~ U(0,0.5), tstop = 10^U(-4,-2).

Physical aspect A maps to the carrier length Lx = 2A (the reference's
``Lx`` is the mirror-doubled periodic grid).
"""
import numpy as np

from .config import ConvectionConfig
from .core import bench_ic
from .basis import shape_ops
from .serve import Request

BENCH_64x96_FIELDS = dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=64, Nz=96, Lx=2.0,
                 precond="bfbt")
C1_TSTOP = 5e-3
HEATED_2TO1_FIELDS = dict(viscosity_model="eta_T", deta_T=1e3, Ra=1.25e4, Nx=64, Nz=96, Lx=4.0,
                 internal_heating=1.5)
HEATED_LOG_TIMES = (0.01, 0.02, 0.03)


class _Grid:
    """The x / z / kx / coeffs of a shape (what the IC helpers need)."""

    def __init__(self, Nx, Nz, Lx):
        o = shape_ops(Nx, Nz, Lx)
        self.x, self.z, self.kx = o.x, o.z, o.kx
        self.Nx, self.V, self.Vi = Nx, o.V, o.Vi

    def coeffs(self, g):
        A = np.fft.rfft(np.asarray(g), axis=0) / self.Nx
        return (self.Vi @ A.T).T

    def seed(self, amp=1e-3, m=1, n=1):
        X, Z = np.meshgrid(self.x, self.z, indexing="ij")
        return self.coeffs(amp * np.cos(m * np.pi * X) * np.sin(n * np.pi * Z))


def bench_64x96_config():
    return ConvectionConfig(**BENCH_64x96_FIELDS)


def bench_64x96_theta0(amp=3e-2):
    g = _Grid(64, 96, 2.0)
    return bench_ic.conductive_theta(g, amp=amp)


def heated_2to1_requests(route="legacy"):
    """The heated 2:1 run on Stokes route ``route`` ('legacy', the
    ConvectionConfig default, or 'bfbt')."""
    cfg = ConvectionConfig(precond=route, **HEATED_2TO1_FIELDS)
    g = _Grid(64, 96, 4.0)
    th = bench_ic.seeded_theta(g, H=1.5, seed=2, sigma=1e-2)
    return [Request("heated-2to1", cfg, th.real, HEATED_LOG_TIMES[-1])]


def _physics_cfg(model, Nx, Nz, Lx, Ra, deta, H, Di=0.0, window=None):
    kw = dict(Nx=int(Nx), Nz=int(Nz), Lx=float(Lx), Ra=float(Ra),
              internal_heating=float(H), dissipation_number=float(Di), precond="bfbt")
    if window is not None:
        kw.update(eta_min=float(window[0]), eta_max=float(window[1]))
    if model == "arrhenius":
        kw.update(viscosity_model="arrhenius", deta_T=float(deta), arrhenius_Ts=1.0,
                  arrhenius_E=2.0 * np.log(float(deta)))
    else:
        kw.update(viscosity_model=model, deta_T=float(deta))
    return ConvectionConfig(**kw)


def ensemble_requests(n=256, seed=3, tstop=1e-3,
                      grids=((32, 24), (64, 48), (64, 96), (128, 96))):
    rng = np.random.default_rng(seed)
    out, cache = [], {}
    for i in range(n):
        Nx, Nz = grids[rng.integers(len(grids))]
        Ra = 10.0 ** rng.uniform(4.0, 6.0)
        deta = 10.0 ** rng.uniform(0.0, 4.0)
        model = ("eta_z", "eta_T", "arrhenius")[rng.integers(3)]
        H = rng.uniform(0.0, 2.0)
        Lx = (2.0, 4.0)[rng.integers(2)]
        ics = int(rng.integers(1 << 31))
        cfg = _physics_cfg(model, Nx, Nz, Lx, Ra, deta, H)
        g = cache.setdefault((Nx, Nz, Lx), _Grid(Nx, Nz, Lx))
        th = bench_ic.seeded_theta(g, H=H, seed=ics, sigma=1e-2)
        out.append(Request("ens%03d" % i, cfg, th.real, tstop, 0.0))
    return out


def stream_requests(n=64, seed=4, rate=1.0,
                    grids=((32, 24), (64, 48), (64, 96), (128, 96), (128, 192), (256, 192)),
                    tstop_range=(-4.0, -2.0), Lx=2.0):
    """``n`` requests with Poisson arrivals of ``rate`` per second."""
    rng = np.random.default_rng(seed)
    out, cache, t = [], {}, 0.0
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        Nx, Nz = grids[rng.integers(len(grids))]
        deta = 10.0 ** rng.uniform(2.0, 5.0)
        Ra = 10.0 ** rng.uniform(4.0, 7.0)
        H = rng.uniform(0.0, 2.0)
        Di = rng.uniform(0.0, 0.5)
        tstop = 10.0 ** rng.uniform(*tstop_range)
        ics = int(rng.integers(1 << 31))
        cfg = _physics_cfg("arrhenius", Nx, Nz, Lx, Ra, deta, H, Di, window=(1e-5, 1.0))
        g = cache.setdefault((Nx, Nz, Lx), _Grid(Nx, Nz, Lx))
        th = bench_ic.seeded_theta(g, H=H, seed=ics, sigma=1e-2)
        out.append(Request("str%03d" % i, cfg, th.real, tstop, t))
    return out
