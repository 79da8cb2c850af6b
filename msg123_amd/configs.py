"""Synthetic problem definitions for the Richards Newton-Krylov step.

Every workload is a structured box of nx*ny*nz cells with uniform spacing,
a closed (no-flow) box except a constant top recharge, and per-cell
van Genuchten-Mualem soil parameters.  All soil arrays are float32, as the
reference stores them (Assign_calculation.f90:19-22), and are generated
deterministically from numpy's PCG64 so that every consumer (the GPU path,
the CPU oracle, the reference deck writer) sees byte-identical inputs.

Units follow the reference's input conventions: lengths in metres, Ks in
m/s, alpha in 1/m, Ss in 1/m, recharge in mm/h (Open_file.f90:744 converts
with 1e-3/3600), times in seconds.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Problem:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    alpha: np.ndarray      # float32 [N] 1/m, natural order (x fastest, then y, then z top->bottom)
    vg_n: np.ndarray       # float32 [N]
    theta_r: np.ndarray    # float32 [N]
    theta_s: np.ndarray    # float32 [N] (porosity)
    ks: np.ndarray         # float32 [N] m/s (isotropic: Kx = Ky = Kz)
    ss: np.ndarray         # float32 [N] 1/m
    wt_depth: float        # initial water-table depth below the top surface, m
    recharge_mmh: float    # top recharge, mm/h (constant in time)
    init_step: float       # s
    max_step: float        # s
    end_time: float        # s (sum of the planned steps)
    maxout_iter: int = 20
    maxinn_iter: int = 1000
    criteria: float = 1e-8  # Newton max-norm head change tolerance, m
    # initial pressure head (m) of every cell instead of a water table at
    # wt_depth (the reference's init_type 6 with a uniform 3D file)
    h0_pressure: float | None = None
    # unit-gradient outflow through the bottom face (no reference
    # counterpart: the reference's bottom is closed)
    free_drainage: bool = False
    # piecewise-constant recharge in time: record k is rech_mmh[k] mm/h,
    # uniform over the top, until rech_etime[k] seconds (float32 each; the
    # last record stays in effect).  None: the constant recharge_mmh.  The
    # reference reads such a series from a recharge file with a time unit
    # (Open_file.f90:744-800, Time_module.f90:274-330,802-858).
    rech_etime: np.ndarray | None = None
    rech_mmh: np.ndarray | None = None

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    def zcoords(self) -> np.ndarray:
        """Float32 node elevations (grid type 7: zcorner + zwidth*(k-1))."""
        k = np.arange(self.nz + 1, dtype=np.float32)
        return np.float32(0.0) + np.float32(-self.dz) * k


def _uniform(rng, lo, hi, n):
    return rng.uniform(lo, hi, n).astype(np.float32)


def make_iid(name, nx, ny, nz, dx, dy, dz, seed, *, wt_depth=5.0,
             recharge_ms=1e-7, dt=3600.0, nsteps=24, lnks_mean=math.log(1e-5),
             lnks_sd=1.0, maxinn_iter=1000, criteria=1e-8) -> Problem:
    """Config-0 style soil: every parameter iid per cell from one seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = nx * ny * nz
    lnks = rng.normal(lnks_mean, lnks_sd, n)
    ks = np.exp(lnks).astype(np.float32)
    alpha = _uniform(rng, 0.5, 3.0, n)
    vg_n = _uniform(rng, 1.3, 2.5, n)
    return Problem(
        name=name, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz,
        alpha=alpha, vg_n=vg_n,
        theta_r=np.full(n, 0.05, np.float32), theta_s=np.full(n, 0.40, np.float32),
        ks=ks, ss=np.full(n, 1e-5, np.float32),
        wt_depth=wt_depth, recharge_mmh=recharge_ms * 1e3 * 3600.0,
        init_step=dt, max_step=dt, end_time=dt * nsteps,
        maxinn_iter=maxinn_iter, criteria=criteria)


def gaussian_random_field(shape, corr, sigma, seed, single=False) -> np.ndarray:
    """Zero-mean GRF with exponential covariance sigma^2 exp(-|r|/corr).

    FFT circulant embedding on a 2x padded periodic grid (Dietrich & Newsam
    1997); negative eigenvalues of the embedding are clipped to zero.  The
    returned field is float64 of the given (nz, ny, nx) shape.  With
    `single`, the embedding runs in float32 with scipy's threaded FFTs (for
    the 134M-cell grid, whose padded embedding has 1.07G points).
    """
    if single:
        return _grf_single(shape, corr, sigma, seed)
    nz, ny, nx = shape
    cz, cy, cx = corr
    pz, py, px = 2 * nz, 2 * ny, 2 * nx
    kz = np.minimum(np.arange(pz), pz - np.arange(pz))[:, None, None] / cz
    ky = np.minimum(np.arange(py), py - np.arange(py))[None, :, None] / cy
    kx = np.minimum(np.arange(px), px - np.arange(px))[None, None, :] / cx
    cov = sigma * sigma * np.exp(-np.sqrt(kz * kz + ky * ky + kx * kx))
    lam = np.fft.rfftn(cov).real
    lam = np.maximum(lam, 0.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    white = rng.standard_normal((pz, py, px))
    field = np.fft.irfftn(np.fft.rfftn(white) * np.sqrt(lam), s=(pz, py, px), axes=(0, 1, 2))
    return field[:nz, :ny, :nx]


def _grf_single(shape, corr, sigma, seed):
    import scipy.fft as sf
    nz, ny, nx = shape
    cz, cy, cx = corr
    pz, py, px = 2 * nz, 2 * ny, 2 * nx
    kz = (np.minimum(np.arange(pz), pz - np.arange(pz)) / cz).astype(np.float32)
    ky = (np.minimum(np.arange(py), py - np.arange(py)) / cy).astype(np.float32)
    kx = (np.minimum(np.arange(px), px - np.arange(px)) / cx).astype(np.float32)
    cov = np.empty((pz, py, px), np.float32)
    for k in range(pz):   # plane by plane keeps the temporaries small
        r = np.sqrt(kz[k] * kz[k] + ky[:, None] ** 2 + kx[None, :] ** 2)
        cov[k] = np.float32(sigma * sigma) * np.exp(-r)
    lam = sf.rfftn(cov, workers=-1).real
    del cov
    np.maximum(lam, 0.0, out=lam)
    np.sqrt(lam, out=lam)
    rng = np.random.Generator(np.random.PCG64(seed))
    white = rng.standard_normal((pz, py, px), dtype=np.float32)
    spec = sf.rfftn(white, workers=-1)
    del white
    spec *= lam
    del lam
    field = sf.irfftn(spec, s=(pz, py, px), workers=-1)
    return field[:nz, :ny, :nx].astype(np.float64)


def make_grf(name, nx, ny, nz, dx, dy, dz, seed, *, corr=(4, 20, 20), sigma=1.5,
             layers_n=(1.4, 1.8, 2.2, 1.6), wt_depth=3.0, recharge_ms=2e-7,
             init_step=600.0, max_step=3600.0, nsteps=48, maxinn_iter=1000,
             criteria=1e-8) -> Problem:
    """Config-1 style soil: lnKs GRF, layered n, iid alpha."""
    n = nx * ny * nz
    f = gaussian_random_field((nz, ny, nx), corr, sigma, seed)
    ks = np.exp(math.log(1e-5) + f).astype(np.float32).reshape(n)
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    alpha = _uniform(rng, 0.5, 3.0, n)
    nl = len(layers_n)
    kl = (np.arange(nz) * nl) // nz
    vg_n = np.repeat(np.asarray(layers_n, np.float32)[kl], nx * ny).astype(np.float32)
    # planned steps: init_step growing x1.2 up to max_step
    dts, dt = [], init_step
    for _ in range(nsteps):
        dts.append(dt)
        dt = min(dt * 1.2, max_step)
    return Problem(
        name=name, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz,
        alpha=alpha, vg_n=vg_n,
        theta_r=np.full(n, 0.05, np.float32), theta_s=np.full(n, 0.40, np.float32),
        ks=ks, ss=np.full(n, 1e-5, np.float32),
        wt_depth=wt_depth, recharge_mmh=recharge_ms * 1e3 * 3600.0,
        init_step=init_step, max_step=max_step, end_time=float(sum(dts)),
        maxinn_iter=maxinn_iter, criteria=criteria)


def config0(seed: int = 0) -> Problem:
    """40x40x20 cells (32k), dx=dy=10 m, dz=0.5 m, iid soil, 24 x 3600 s."""
    return make_iid("c0", 40, 40, 20, 10.0, 10.0, 0.5, seed)


def config1(seed: int = 1, nsteps: int = 48) -> Problem:
    """256x256x64 cells (4.2M), dx=dy=5 m, dz=0.25 m, GRF lnKs, 4 layers."""
    return make_grf("c1", 256, 256, 64, 5.0, 5.0, 0.25, seed, nsteps=nsteps)


def config2(seed: int = 2, ndays: int = 365) -> Problem:
    """Config-1 grid and soil statistics (seed 2) under seasonal recharge
    1e-7 (1 + 0.8 sin(2 pi t / 365 d)) m/s, as daily records (each day's rate
    at the day's start, in mm/h), for `ndays` days; steps from 600 s growing
    x1.2 up to one day, cut at the day ends by the reference's time
    control."""
    p = make_grf("c2", 256, 256, 64, 5.0, 5.0, 0.25, seed, init_step=600.0,
                 max_step=86400.0, nsteps=1)
    day = 86400.0
    k = np.arange(ndays)
    rate = 1e-7 * (1.0 + 0.8 * np.sin(2.0 * np.pi * k / 365.0))
    p.rech_etime = ((k + 1) * day).astype(np.float32)
    p.rech_mmh = (rate * 1e3 * 3600.0).astype(np.float32)
    p.recharge_mmh = float(p.rech_mmh[0])
    p.end_time = float(ndays * day)
    return p


def config4(seed: int = 4, nsteps: int = 12, dims=(1024, 1024, 128)) -> Problem:
    """1024x1024x128 cells (134M), dx=dy=5 m, dz=0.2 m: lnKs GRF with
    correlation 40x40x6 cells and sigma 2, n ~ U[1.5, 3] and alpha ~ U[0.5, 3]
    iid, a pressure head of -10 m everywhere (`h0_pressure`), infiltration
    5e-6 m/s (18 mm/h), 12 steps of 600 s, free-drainage bottom (an
    extension: the reference's bottom is closed).  `dims` = (nx, ny, nz)
    builds the same statistics on a smaller box."""
    nx, ny, nz = dims
    n = nx * ny * nz
    f = gaussian_random_field((nz, ny, nx), (6, 40, 40), 2.0, seed, single=True)
    ks = np.exp(math.log(1e-5) + f).astype(np.float32).reshape(n)
    del f
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    alpha = _uniform(rng, 0.5, 3.0, n)
    vg_n = _uniform(rng, 1.5, 3.0, n)
    p = Problem(
        name="c4", nx=nx, ny=ny, nz=nz, dx=5.0, dy=5.0, dz=0.2,
        alpha=alpha, vg_n=vg_n,
        theta_r=np.full(n, 0.05, np.float32), theta_s=np.full(n, 0.40, np.float32),
        ks=ks, ss=np.full(n, 1e-5, np.float32),
        wt_depth=10.0, recharge_mmh=5e-6 * 1e3 * 3600.0,
        init_step=600.0, max_step=600.0, end_time=600.0 * nsteps, h0_pressure=-10.0,
        free_drainage=True)
    return p


def config3(count: int = 64, nsteps: int = 24) -> list:
    """64 independent hillslope boxes of uneven size: instance i draws from
    seed 1000 + i nx, ny ~ U{32..128}, nz ~ U{16..48}, a water-table depth
    ~ U[2, 8] m and a recharge ~ U[5e-8, 5e-7] m/s; soil as config 0 (iid
    lnKs ~ N(ln 1e-5, 1), alpha ~ U[0.5, 3], n ~ U[1.3, 2.5]); dx=dy=10 m,
    dz=0.5 m; 24 hourly steps."""
    out = []
    for i in range(count):
        rng = np.random.Generator(np.random.PCG64(1000 + i))
        nx, ny = (int(v) for v in rng.integers(32, 129, 2))
        nz = int(rng.integers(16, 49))
        wt = float(rng.uniform(2.0, 8.0))
        rech = float(rng.uniform(5e-8, 5e-7))
        out.append(make_iid(f"c3_{i}", nx, ny, nz, 10.0, 10.0, 0.5, 1000 + i, wt_depth=wt,
                            recharge_ms=rech, nsteps=nsteps))
    return out


def small(nx=6, ny=5, nz=4, seed=3, nsteps=3, dt=3600.0, **kw) -> Problem:
    """Tiny config-0 style box for unit tests."""
    return make_iid(f"s{nx}x{ny}x{nz}", nx, ny, nz, 10.0, 10.0, 0.5, seed,
                    dt=dt, nsteps=nsteps, **kw)
