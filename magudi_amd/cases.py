"""Problem set-ups for the magudi RHS hot path (host side, numpy only).

A ``Case`` is the in-memory equivalent of what magudi reads from ``magudi.inp``,
``bc.dat`` and the PLOT3D grid / function files of an example directory
(``examples/AcousticMonopole/{magudi.inp,bc.dat,config.py}`` in the magudi tree):
grid coordinates, solver options, the boundary-condition patch table, target
state, control/target mollifiers and acoustic sources.  The generators here are
written from the problem statements (uniform Cartesian grids, quiescent
state, Gaussian monopole, tanh-edged mollifiers); they do not use magudi's
Python utilities.

Layout of every field is magudi's ``x(nGridPoints, nComponents)``: a
C-contiguous numpy array of shape ``(nComponents, nz, ny, nx)`` flattened to
``(nComponents, N)`` with the x index fastest.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

# Patch types (bc.dat "Type" column) -> integer codes used across the C ABI.
FAR_FIELD = 1      # SAT_FAR_FIELD
SPONGE = 2         # SPONGE
ACTUATOR = 3       # ACTUATOR
COST_TARGET = 4    # COST_TARGET
KOLMOGOROV = 5     # KOLMOGOROV_FORCING
IMPENETRABLE_WALL = 6   # SAT_SLIP_WALL
ISOTHERMAL_WALL = 7     # SAT_ISOTHERMAL_WALL

PATCH_TYPE_NAMES = {
    "SAT_FAR_FIELD": FAR_FIELD,
    "SPONGE": SPONGE,
    "ACTUATOR": ACTUATOR,
    "COST_TARGET": COST_TARGET,
    "KOLMOGOROV_FORCING": KOLMOGOROV,
    "SAT_SLIP_WALL": IMPENETRABLE_WALL,
    "SAT_ISOTHERMAL_WALL": ISOTHERMAL_WALL,
}

NONE, PLANE, OVERLAP = 0, 1, 2      # grid periodicity types (Grid_enum)
SBP_4_8 = 0


@dataclasses.dataclass
class Patch:
    """One row of bc.dat plus the patch options read from magudi.inp."""

    name: str
    type: int
    normal_dir: int
    extent: tuple          # (iMin, iMax, jMin, jMax, kMin, kMax), 1-based, <0 from end
    params: tuple = (0.0, 0.0, 0.0, 0.0)


@dataclasses.dataclass
class Source:
    location: tuple
    amplitude: float
    frequency: float
    radius: float
    phase: float = 0.0


@dataclasses.dataclass
class Case:
    name: str
    nd: int
    n: tuple                       # grid size (nx, ny, nz)
    coords: np.ndarray             # (nd, N)
    periodicity: tuple = (NONE, NONE, NONE)
    periodic_length: tuple = (0.0, 0.0, 0.0)
    scheme: int = SBP_4_8
    viscous: bool = False
    dissipation: bool = True
    composite_dissipation: bool = False
    gamma: float = 1.4
    reynolds: float = 0.0
    prandtl: float = 0.72
    power_law: float = 0.0
    bulk_ratio: float = 0.0
    dissipation_amount: float = 0.005
    patches: List[Patch] = dataclasses.field(default_factory=list)
    target: Optional[np.ndarray] = None          # (nv, N)
    control_mollifier: Optional[np.ndarray] = None
    target_mollifier: Optional[np.ndarray] = None
    sources: List[Source] = dataclasses.field(default_factory=list)
    dt: float = 0.0
    nsteps: int = 0
    save_interval: int = 0
    initial: Optional[np.ndarray] = None         # (nv, N); defaults to target

    @property
    def nv(self) -> int:
        return self.nd + 2

    @property
    def npoints(self) -> int:
        return int(np.prod(self.n))

    def initial_state(self) -> np.ndarray:
        q = self.initial if self.initial is not None else self.target
        return np.ascontiguousarray(q, dtype=np.float64).copy()

    def patch_points(self, p: Patch) -> np.ndarray:
        """Flat grid indices of a patch in the reference's patch order (i fastest)."""
        lo, hi = [], []
        for d in range(3):
            a, b = p.extent[2 * d], p.extent[2 * d + 1]
            if d >= self.nd:
                a, b = 1, 1
            if a < 0:
                a += self.n[d] + 1
            if b < 0:
                b += self.n[d] + 1
            lo.append(a - 1)
            hi.append(b - 1)
        k, j, i = np.meshgrid(np.arange(lo[2], hi[2] + 1), np.arange(lo[1], hi[1] + 1),
                              np.arange(lo[0], hi[0] + 1), indexing="ij")
        return (i + self.n[0] * (j + self.n[1] * k)).ravel()

    def patches_of(self, ptype: int) -> List[Patch]:
        return [p for p in self.patches if p.type == ptype]


def quiescent_state(nd: int, npts: int, gamma: float = 1.4) -> np.ndarray:
    """rho = 1, u = 0, p = 1/gamma (speed of sound 1)."""
    q = np.zeros((nd + 2, npts))
    q[0] = 1.0
    q[nd + 1] = (1.0 / gamma) / (gamma - 1.0)
    return q


def _tanh_window(r, r0, r1, sharp):
    return 0.25 * (1.0 + np.tanh(sharp * (r - r0))) * (1.0 - np.tanh(sharp * (r - r1)))


def _box_extent(x1d, lo, hi):
    """1-based inclusive index range of grid points with lo <= x <= hi."""
    idx = np.nonzero((x1d >= lo) & (x1d <= hi))[0]
    return int(idx[0]) + 1, int(idx[-1]) + 1


def acoustic_monopole(n: int = 201, nsteps: int = 480, save_interval: int = 32,
                      with_control: bool = True, composite: bool = False) -> Case:
    """2-D acoustic monopole on [-14, 14]^2 (inviscid Euler, SBP 4-8).

    Gaussian monopole at the origin (amplitude 1e-2, frequency 1/30, radius 1.5
    i.e. width 0.5 in exp(-9 r^2 / (2 R^2))); far-field SAT and 10 %-width
    sponges on all sides; cost target on the annulus 7 <= r <= 9; control
    mollifier disk of radius 1 at (0, 4).  dt = 0.5 dx / sqrt(2) (CFL 0.5, c = 1).
    """
    x1 = np.linspace(-14.0, 14.0, n)
    X, Y = np.meshgrid(x1, x1, indexing="xy")           # X varies along axis 1 (i)
    coords = np.stack([X.ravel(), Y.ravel()])
    npts = n * n
    gamma = 1.4
    w = n // 10 + 1
    patches = [
        Patch("farField.E", FAR_FIELD, 1, (1, 1, 1, -1, 1, -1), (1.0, 0.0, 0, 0)),
        Patch("farField.W", FAR_FIELD, -1, (-1, -1, 1, -1, 1, -1), (1.0, 0.0, 0, 0)),
        Patch("farField.S", FAR_FIELD, 2, (1, -1, 1, 1, 1, -1), (1.0, 0.0, 0, 0)),
        Patch("farField.N", FAR_FIELD, -2, (1, -1, -1, -1, 1, -1), (1.0, 0.0, 0, 0)),
        Patch("sponge.E", SPONGE, 1, (1, w, 1, -1, 1, -1), (0.2, 2.0, 0, 0)),
        Patch("sponge.W", SPONGE, -1, (-w, -1, 1, -1, 1, -1), (0.2, 2.0, 0, 0)),
        Patch("sponge.S", SPONGE, 2, (1, -1, 1, w, 1, -1), (0.2, 2.0, 0, 0)),
        Patch("sponge.N", SPONGE, -2, (1, -1, -w, -1, 1, -1), (0.2, 2.0, 0, 0)),
    ]
    r = np.hypot(coords[0], coords[1])
    tgt_moll = _tanh_window(r, 7.0, 9.0, 4.0)
    ti = _box_extent(x1, -9.6, 9.6)
    patches.append(Patch("targetRegion", COST_TARGET, 0, (ti[0], ti[1], ti[0], ti[1], 1, -1)))
    ctrl_moll = None
    if with_control:
        rc = np.hypot(coords[0] - 0.0, coords[1] - 4.0)
        ctrl_moll = 0.5 * (1.0 - np.tanh(6.0 * (rc - 1.0)))
        ci = _box_extent(x1, -1.6, 1.6)
        cj = _box_extent(x1, 2.4, 5.6)
        patches.append(Patch("controlRegion", ACTUATOR, 0, (ci[0], ci[1], cj[0], cj[1], 1, -1)))
    dx = 28.0 / (n - 1)
    dt = 0.5 * dx / math.sqrt(2.0)
    case = Case(
        name=f"monopole{n}", nd=2, n=(n, n, 1), coords=coords, gamma=gamma,
        dissipation=True, composite_dissipation=composite, dissipation_amount=0.005,
        patches=patches, target=quiescent_state(2, npts, gamma),
        control_mollifier=ctrl_moll if ctrl_moll is not None else np.zeros(npts),
        target_mollifier=tgt_moll,
        sources=[Source((0.0, 0.0, 0.0), 1e-2, 1.0 / 30.0, 1.5)],
        dt=dt, nsteps=nsteps, save_interval=save_interval)
    return case


def kolmogorov3d(n: int = 256, nsteps: int = 200, seed: int = 20261202,
                 wavenumber: int = 4, amplitude: float = 0.01, mach: float = 0.2,
                 reynolds: float = 100.0, prandtl: float = 0.7, rms: float = 0.05) -> Case:
    """3-D viscous periodic Kolmogorov-type flow on [0, 1)^3 (PLANE periodicity).

    Forcing f_x = A sin(2 pi n y) with integer n (the only form magudi's
    KOLMOGOROV_FORCING patch expresses); initial velocity is a seeded
    random-phase solenoidal field with spectrum k^4 exp(-2 (k/4)^2) scaled to the
    given rms.  Reynolds number is magudi's acoustic Reynolds number Re / Mach.
    """
    x1 = np.arange(n) / n
    Z, Y, X = np.meshgrid(x1, x1, x1, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    npts = n ** 3
    gamma = 1.4
    q = quiescent_state(3, npts, gamma)
    u = solenoidal_field(n, seed, rms)
    q[1:4] = u
    q[4] = q[4] + 0.5 * (u ** 2).sum(axis=0)
    patches = [Patch("forcing", KOLMOGOROV, 0, (1, -1, 1, -1, 1, -1),
                     (float(wavenumber), amplitude, 0, 0))]
    dx = 1.0 / n
    dt = 0.5 * dx / (1.0 + 3 * rms) / math.sqrt(3.0)
    return Case(
        name=f"kolmogorov{n}", nd=3, n=(n, n, n), coords=coords,
        periodicity=(PLANE, PLANE, PLANE), periodic_length=(1.0, 1.0, 1.0),
        viscous=True, dissipation=False, gamma=gamma, reynolds=reynolds / mach,
        prandtl=prandtl, power_law=0.0, bulk_ratio=0.0, patches=patches,
        target=quiescent_state(3, npts, gamma), control_mollifier=np.zeros(npts),
        target_mollifier=np.zeros(npts), dt=dt, nsteps=nsteps, save_interval=max(1, nsteps // 8),
        initial=q)


def solenoidal_field(n: int, seed: int, rms: float) -> np.ndarray:
    """Random-phase divergence-free velocity, E(k) ~ k^4 exp(-2 (k/4)^2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    amp = np.where(kk > 0, np.sqrt(kk ** 4 * np.exp(-2.0 * (kk / 4.0) ** 2)) / np.maximum(kk, 1), 0.0)
    comps = []
    for _ in range(3):
        phase = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
        comps.append(amp * np.exp(1j * phase))
    uh = np.stack(comps)
    kvec = np.stack([KX, KY, KZ])
    kdotu = (kvec * uh).sum(axis=0)
    uh = uh - kvec * kdotu / np.where(kk > 0, kk ** 2, 1.0)
    u = np.real(np.fft.ifftn(uh, axes=(1, 2, 3)))
    u *= rms / np.sqrt((u ** 2).sum(axis=0).mean())
    return u.reshape(3, -1)


def walled_monopole(n: int = 41, viscous: bool = False, nsteps: int = 8,
                    save_interval: int = 4) -> Case:
    """Monopole with walls on the y faces: a slip (impenetrable) wall at the bottom and,
    for viscous flow, an isothermal no-slip wall at the top (else a slip wall); far
    field and sponges on the x faces.  Wall patch options: (inviscid_penalty_amount,
    viscous_penalty_amount, wall temperature)."""
    c = acoustic_monopole(n=n, nsteps=nsteps, save_interval=save_interval)
    tw = 1.0 / (c.gamma - 1.0)
    keep = [p for p in c.patches if p.normal_dir not in (2, -2)]
    top = ISOTHERMAL_WALL if viscous else IMPENETRABLE_WALL
    keep[0:0] = [Patch("wall.S", IMPENETRABLE_WALL, 2, (1, -1, 1, 1, 1, -1), (1.0, 0.0, 0, 0)),
                 Patch("wall.N", top, -2, (1, -1, -1, -1, 1, -1), (1.0, 1.0, tw, 0))]
    c.patches = keep
    c.name = f"walled_monopole{n}" + ("_viscous" if viscous else "")
    if viscous:
        c.viscous = True
        c.reynolds = 200.0
        c.prandtl = 0.7
    return c


def periodic_wave_1d(n: int = 2048, nsteps: int = 64, save_interval: int = 16,
                     amplitude: float = 1e-2) -> Case:
    """1-D periodic acoustic wave on [0, 1) (PLANE periodicity): a right-running
    Gaussian pressure pulse (p' = rho' = amplitude g, u' = p'), SBP 4-8 with
    dissipation; cost target on [0.6, 0.8], actuator on [0.2, 0.3]."""
    x = np.arange(n) / n
    coords = x[None, :].copy()
    gamma = 1.4
    q = quiescent_state(1, n, gamma)
    g = np.exp(-((x - 0.5) / 0.05) ** 2)
    dp = amplitude * g
    q[0] = q[0] + dp
    q[1] = q[0] * dp
    q[2] = (1.0 / gamma + dp) / (gamma - 1.0) + 0.5 * q[1] ** 2 / q[0]
    ti = _box_extent(x, 0.6, 0.8)
    ci = _box_extent(x, 0.2, 0.3)
    patches = [Patch("targetRegion", COST_TARGET, 0, (ti[0], ti[1], 1, -1, 1, -1)),
               Patch("controlRegion", ACTUATOR, 0, (ci[0], ci[1], 1, -1, 1, -1))]
    dx = 1.0 / n
    return Case(
        name=f"wave1d_{n}", nd=1, n=(n, 1, 1), coords=coords, periodicity=(PLANE, NONE, NONE),
        periodic_length=(1.0, 0.0, 0.0), gamma=gamma, dissipation=True, dissipation_amount=0.005,
        patches=patches, target=quiescent_state(1, n, gamma),
        control_mollifier=_tanh_window(x, 0.2, 0.3, 200.0),
        target_mollifier=_tanh_window(x, 0.6, 0.8, 200.0),
        dt=0.5 * dx, nsteps=nsteps, save_interval=save_interval, initial=q)


def mixing_layer(nx: int = 513, ny: int = 257, stretch: float = 1.02, velocity_ratio: float = 0.5,
                 thickness: float = 1.0, perturbation: float = 1e-3, seed: int = 0,
                 nsteps: int = 64, save_interval: int = 16) -> Case:
    """2-D temporally developing mixing layer: periodic in x (PLANE), y stretched
    geometrically away from the centre line (spacing ratio `stretch`), tanh velocity
    profile u = (U1 + U2)/2 + (U1 - U2)/2 tanh(2 y / thickness) with U2 / U1 =
    velocity_ratio and U1 = 0.5 (Mach 0.5), a seeded random perturbation of the
    velocity; far field and sponges on the y faces; cost target on a band above the
    layer, actuator on a patch in it.  Inviscid Euler with dissipation."""
    ly_half_cells = (ny - 1) // 2
    dy0 = 0.2 * thickness
    ys = np.concatenate([[0.0], np.cumsum(dy0 * stretch ** np.arange(ly_half_cells))])
    y1 = np.concatenate([-ys[::-1], ys[1:]]) if (ny - 1) % 2 == 0 else np.concatenate([-ys[::-1], ys[1:], [ys[-1] + dy0]])
    y1 = y1[:ny]
    lx = 30.0 * thickness
    x1 = np.arange(nx) * (lx / nx)
    X, Y = np.meshgrid(x1, y1, indexing="xy")
    coords = np.stack([X.ravel(), Y.ravel()])
    npts = nx * ny
    gamma = 1.4
    u1 = 0.5
    u2 = velocity_ratio * u1
    U = 0.5 * (u1 + u2) + 0.5 * (u1 - u2) * np.tanh(2.0 * coords[1] / thickness)
    rng = np.random.Generator(np.random.PCG64(seed))
    env = np.exp(-(coords[1] / (2.0 * thickness)) ** 2)
    du = perturbation * env * rng.standard_normal(npts)
    dv = perturbation * env * rng.standard_normal(npts)
    target = quiescent_state(2, npts, gamma)
    target[1] = U
    target[3] = target[3] + 0.5 * U ** 2
    q = target.copy()
    q[1] = U + du
    q[2] = dv
    q[3] = (1.0 / gamma) / (gamma - 1.0) + 0.5 * ((U + du) ** 2 + dv ** 2)
    w = max(12, ny // 10)
    patches = [
        Patch("farField.S", FAR_FIELD, 2, (1, -1, 1, 1, 1, -1), (1.0, 0.0, 0, 0)),
        Patch("farField.N", FAR_FIELD, -2, (1, -1, -1, -1, 1, -1), (1.0, 0.0, 0, 0)),
        Patch("sponge.S", SPONGE, 2, (1, -1, 1, w, 1, -1), (0.2, 2.0, 0, 0)),
        Patch("sponge.N", SPONGE, -2, (1, -1, -w, -1, 1, -1), (0.2, 2.0, 0, 0)),
    ]
    tj = _box_extent(y1, 2.0 * thickness, 6.0 * thickness)
    patches.append(Patch("targetRegion", COST_TARGET, 0, (1, -1, tj[0], tj[1], 1, -1)))
    ci = _box_extent(x1, 0.2 * lx, 0.3 * lx)
    cj = _box_extent(y1, -1.0 * thickness, 1.0 * thickness)
    patches.append(Patch("controlRegion", ACTUATOR, 0, (ci[0], ci[1], cj[0], cj[1], 1, -1)))
    tgt_moll = _tanh_window(coords[1], 2.0 * thickness, 6.0 * thickness, 4.0)
    ctrl_moll = _tanh_window(coords[0], 0.2 * lx, 0.3 * lx, 4.0) * _tanh_window(coords[1], -thickness, thickness, 4.0)
    dxmin = min(lx / nx, dy0)
    dt = 0.5 * dxmin / (1.0 + u1)
    return Case(
        name=f"mixing_layer{nx}x{ny}", nd=2, n=(nx, ny, 1), coords=coords,
        periodicity=(PLANE, NONE, NONE), periodic_length=(lx, 0.0, 0.0), gamma=gamma,
        dissipation=True, dissipation_amount=0.005, patches=patches, target=target,
        control_mollifier=ctrl_moll, target_mollifier=tgt_moll, dt=dt, nsteps=nsteps,
        save_interval=save_interval, initial=q)


def uneven_batch(seed: int = 5, n_waves: int = 8, n_monopoles: int = 16, n_layers: int = 2,
                 scale: float = 1.0, nsteps: int = 32, save_interval: int = 8) -> List[Case]:
    """The mixed batch of problem types: 1-D periodic waves, 2-D monopoles on odd grid
    sizes drawn from [201, 513] and mixing layers; `scale` shrinks the grids for tests."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out: List[Case] = []
    for k in range(n_waves):
        out.append(periodic_wave_1d(n=max(64, int(2048 * scale)), nsteps=nsteps, save_interval=save_interval))
    for k in range(n_monopoles):
        n = int(rng.integers(100, 257)) * 2 + 1
        n = max(41, int(n * scale) | 1)
        out.append(acoustic_monopole(n=n, nsteps=nsteps, save_interval=save_interval))
    for k in range(n_layers):
        out.append(mixing_layer(nx=max(64, int(513 * scale)), ny=max(41, int(257 * scale) | 1),
                                seed=k, nsteps=nsteps, save_interval=save_interval))
    return out


def small_monopole(n: int = 41, nsteps: int = 8, save_interval: int = 4) -> Case:
    """Scaled-down monopole used by the unit and parity tests."""
    c = acoustic_monopole(n=n, nsteps=nsteps, save_interval=save_interval)
    return c


def random_state(case: Case, seed: int, amplitude: float = 1e-2) -> np.ndarray:
    """Target state plus a smooth seeded perturbation (keeps rho, T positive)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    q = case.target.copy()
    nd = case.nd
    x = case.coords
    pert = np.zeros_like(q)
    for l in range(case.nv):
        phase = rng.uniform(0, 2 * np.pi, size=nd)
        kk = rng.uniform(0.2, 1.0, size=nd)
        f = np.ones(case.npoints)
        for d in range(nd):
            f = f * np.cos(kk[d] * x[d] * 2 * np.pi / max(np.ptp(x[d]), 1e-30) * 3 + phase[d])
        pert[l] = f
    q = q + amplitude * pert
    q[0] = np.abs(q[0]) + 0.5
    return q
