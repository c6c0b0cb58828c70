"""Input decks of the benchmark and test configurations.

``Deck`` mirrors the parameter block of the reference's ``_Input_Deck``
(src/input/deck.h:161-348): the same member names and defaults, the same
``courant_length`` and ``derive_params``, and ``driver_constants`` evaluates
the constants of example/example.cpp:77-114,179-181 with the reference's
real_t arithmetic (numpy scalars of the deck's dtype; double literals are
applied in float64 exactly where the C++ source promotes).

Particle loading for the synthetic configurations goes through the
library's counter-based plasma loader (``cpic_init_plasma``), so the C++
driver and Python produce bit-identical initial states.
"""
import ctypes
import math

import numpy as np

from . import _capi

F64 = np.float64


class Deck:
    def __init__(self, dtype=np.float32):
        R = self.R = np.dtype(dtype).type
        # _Input_Deck member defaults (deck.h:254-309)
        self.de = R(1.0)
        self.ec = R(1.0)
        self.me = R(1.0)
        self.mu = R(1.0)
        self.c = R(1.0)
        self.eps = R(1.0)
        self.qsp = R(-self.ec)
        self.n0 = R(1.0)
        self.nx, self.ny, self.nz = 16, 1, 1
        self.num_ghosts = 1
        self.nppc = 1
        self.dt = R(1.0)
        self.num_steps = 2
        self.len_x_global = R(1.0)
        self.len_y_global = R(1.0)
        self.len_z_global = R(1.0)
        self.Npe = R(-1)
        self.Ne = R(-1)
        self.v0 = R(1.0)
        self.num_particles = -1
        self.plasma = None  # cpic_plasma_spec fields for synthetic decks
        self.name = "default"

    # deck.h:191-198
    def courant_length(self, lx, ly, lz, nx, ny, nz):
        R = self.R
        w1 = R(0)
        for n, l in ((nx, lx), (ny, ly), (nz, lz)):
            if n > 1:
                w0 = R(R(n) / R(l))
                w1 = R(w1 + R(w0 * w0))
        # sqrt of a real_t: double rounding of a correctly rounded sqrt is
        # innocuous, so float64 sqrt then rounding equals the float sqrt
        return R(math.sqrt(F64(R(R(1) / w1))))

    # deck.h:323-348
    def derive_params(self):
        R = self.R
        self.len_x, self.len_y, self.len_z = self.len_x_global, self.len_y_global, self.len_z_global
        self.dx = R(self.len_x / R(self.nx))
        self.dy = R(self.len_y / R(self.ny))
        self.dz = R(self.len_z / R(self.nz))
        g = 2 * self.num_ghosts
        self.num_cells = (self.nx + g) * (self.ny + g) * (self.nz + g)
        if self.num_particles < 0:
            self.num_particles = self.nx * self.ny * self.nz * self.nppc
            if self.Ne < 0:
                self.Ne = R(self.num_particles)
        if self.Npe < 0:
            self.Npe = R(R(R(self.n0 * self.len_x_global) * self.len_y_global) * self.len_z_global)
        return self

    # example/example.cpp:77-114,179-181
    def driver_constants(self):
        R = self.R
        c, dt = self.c, self.dt
        k = {}
        k["dxp"] = R(np.float32(2.0) / np.float32(self.nppc))
        k["qdt_2mc"] = R(R(self.qsp * dt) / R(R(R(2) * self.me) * c))
        k["cdt_dx"] = R(R(c * dt) / self.dx)
        k["cdt_dy"] = R(R(c * dt) / self.dy)
        k["cdt_dz"] = R(R(c * dt) / self.dz)
        k["dt_eps0"] = R(dt / self.eps)
        frac = R(1.0)
        k["we"] = R(R(self.Npe) / R(int(self.Ne)))
        k["px"] = R(R(R(frac * c) * dt) / self.dx) if self.nx > 1 else R(0)
        k["py"] = R(R(R(frac * c) * dt) / self.dy) if self.ny > 1 else R(0)
        k["pz"] = R(R(R(frac * c) * dt) / self.dz) if self.nz > 1 else R(0)
        k.update(qsp=self.qsp, dx=self.dx, dy=self.dy, dz=self.dz, dt=dt, me=self.me, c=c)
        return k

    def step_params(self, **kw):
        k = self.driver_constants()
        names = ["qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp", "dx", "dy", "dz", "dt",
                 "px", "py", "pz", "dt_eps0", "me", "c"]
        d = {n: float(k[n]) for n in names}
        d.update(kw)
        return d

    def load_particles(self):
        """Initial particles of a synthetic deck (SoA dict + int32 cell)."""
        if self.plasma is None:
            raise ValueError(f"deck {self.name} has no synthetic plasma")
        n = self.num_particles
        out = {k: np.empty(n, self.R) for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
        out["cell"] = np.empty(n, np.int32)
        spec = _capi.PlasmaSpec()
        p = self.plasma
        spec.seed = p["seed"]
        spec.nx, spec.ny, spec.nz, spec.nppc = self.nx, self.ny, self.nz, self.nppc
        spec.weight = float(self.driver_constants()["we"])
        for i in range(3):
            spec.uth[i] = p.get("uth", (0, 0, 0))[i]
            spec.drift[i] = p.get("drift", (0, 0, 0))[i]
        spec.pert_amp = p.get("pert_amp", 0.0)
        spec.len_x = float(self.len_x)
        spec.pert_mode = p.get("pert_mode", 0)
        spec.counter_streaming = 1 if p.get("counter_streaming") else 0
        self._column_spec(spec)
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _capi.check(_capi.load().cpic_init_plasma(
            ctypes.byref(spec), np.dtype(self.R).itemsize, n,
            *[ptr(out[k]) for k in ("x", "y", "z", "ux", "uy", "uz", "w")], ptr(out["cell"])))
        return out

    def _column_spec(self, spec):
        p = self.plasma
        if p.get("profile", 0) == 1:
            spec.profile = 1
            spec.len_y = float(self.len_y)
            spec.ring_radius = p["ring_radius"]
            spec.ring_width = p["ring_width"]
            spec.column_charge = p["column_charge"]
            spec.guide_bz = p["guide_bz"]
            spec.pert_amp = p.get("pert_amp", 0.0)
            spec.pert_mode = p.get("pert_mode", 0)

    def load_fields(self):
        """Initial fields ([9][Ng], reference member order) of a deck whose
        plasma is an electron column: its radial E on the Yee edges of every
        voxel and the guide field in cbz.  None for decks that start from
        zero fields."""
        if self.plasma is None or self.plasma.get("profile", 0) != 1:
            return None
        spec = _capi.PlasmaSpec()
        spec.nx, spec.ny, spec.nz = self.nx, self.ny, self.nz
        spec.len_x = float(self.len_x)
        self._column_spec(spec)
        f = np.empty((9, self.num_cells), self.R)
        _capi.check(_capi.load().cpic_init_column_fields(
            ctypes.byref(spec), np.dtype(self.R).itemsize, f.ctypes.data_as(ctypes.c_void_p)))
        return f


def _cube(deck, n, nppc, dx, steps):
    R = deck.R
    deck.nx = deck.ny = deck.nz = n
    deck.nppc = nppc
    deck.num_steps = steps
    L = R(F64(n) * F64(dx))
    deck.len_x_global = deck.len_y_global = deck.len_z_global = L
    deck.dt = R(F64(0.99) * F64(deck.courant_length(L, L, L, n, n, n)) / F64(deck.c))
    return deck


def weibel_3d(dtype=np.float32, n=128, nppc=32, steps=500, seed=11):
    """Precision/metric configuration: 128^3 periodic cells, 32 ppc,
    bi-Maxwellian electrons u_th = (0.3, 0.05, 0.05) c, immobile ions,
    dx = 0.1 c/w_pe, dt = 0.99 Courant."""
    d = _cube(Deck(dtype), n, nppc, 0.1, steps)
    d.name = f"weibel_{n}^3x{nppc}"
    d.n0 = d.R(1.0)
    d.plasma = dict(seed=seed, uth=(0.3, 0.05, 0.05))
    return d.derive_params()


def two_stream_3d(dtype=np.float32, n=64, nppc=64, steps=200, seed=7, v0=0.2, uth=0.01,
                  amp=1e-3, length=None):
    """3D two-stream: counter-streaming beams +-v0 along x (nppc/2 each),
    Maxwellian spread uth per component, mode-1 drift perturbation.  The box
    edge is `length`, by default 2 pi v0 / 0.5 (derivatives with respect to
    v0 hold it fixed)."""
    d = Deck(dtype)
    R = d.R
    d.nx = d.ny = d.nz = n
    d.nppc = nppc
    d.num_steps = steps
    d.v0 = R(v0)
    # fastest-growing two-stream wavelength scale
    L = R(2.0 * math.pi * v0 / 0.5) if length is None else R(length)
    d.len_x_global = d.len_y_global = d.len_z_global = L
    d.dt = R(F64(0.99) * F64(d.courant_length(L, L, L, n, n, n)) / F64(d.c))
    d.n0 = R(1.0)
    u0 = v0 / math.sqrt(1.0 - v0 * v0)
    d.name = f"two_stream_{n}^3x{nppc}"
    d.plasma = dict(seed=seed, uth=(uth, uth, uth), drift=(u0, 0.0, 0.0), pert_amp=amp,
                    pert_mode=1, counter_streaming=True)
    return d.derive_params()


def two_stream_1d(dtype=np.float64, nx=32, nppc_per_beam=32, steps=1000, seed=1234, v0=0.2,
                  amp=1e-3):
    """1D-periodic two-stream in x (the reference's 2stream-short case with
    the beams spread along x): nx x 1 x 1 cells, Lx = 2 pi 0.2 / 0.5, two
    cold counter-streaming electron beams at +-v0 c along x (nppc_per_beam
    each, alternating), offsets uniform-random from the seed, drift
    modulated by 1 + amp sin(2 pi x / Lx), immobile ions, EM solver,
    dt = 0.99 Courant."""
    d = Deck(dtype)
    R = d.R
    d.nx, d.ny, d.nz = nx, 1, 1
    d.nppc = 2 * nppc_per_beam
    d.num_steps = steps
    d.v0 = R(v0)
    d.len_x_global = R(2.0 * math.pi * v0 / 0.5)
    d.dt = R(F64(0.99) * F64(d.courant_length(d.len_x_global, d.len_y_global, d.len_z_global,
                                                nx, 1, 1)) / F64(d.c))
    d.n0 = R(1.0)
    d.name = f"two_stream_1d_{nx}x{d.nppc}"
    d.plasma = dict(seed=seed, drift=(v0 / math.sqrt(1.0 - v0 * v0), 0.0, 0.0), pert_amp=amp,
                    pert_mode=1, counter_streaming=True)
    return d.derive_params()


def dioctron_3d(dtype=np.float32, nx=192, ny=192, nz=64, nppc=32, steps=1000, seed=3, bz=1.0,
                amp=1e-3, mode=3, dx=0.1):
    """Magnetized electron column (a 3D version of the reference's
    decks/dioctron_3d.cxx): nx x ny x nz periodic cells of edge dx, nppc
    electrons in every cell whose weights carry the ring density
    n(r) = n0 exp(-((r - 0.3 Lx) / 0.05 Lx)^2) (1 + amp cos(mode theta)) about
    the z axis through the box centre, n0 = 0.25 so that with the guide field
    bz = 1 the peak w_ce / w_pe = 2; each electron drifts at the E x B
    velocity of the column's own radial field, which also initializes ex, ey
    in every voxel (ghosts included, as the reference's deck does); immobile
    neutralizing ions are not part of the field solve; dt = 0.99 Courant."""
    d = Deck(dtype)
    R = d.R
    d.nx, d.ny, d.nz = nx, ny, nz
    d.nppc = nppc
    d.num_steps = steps
    d.len_x_global = R(F64(nx) * F64(dx))
    d.len_y_global = R(F64(ny) * F64(dx))
    d.len_z_global = R(F64(nz) * F64(dx))
    d.dt = R(F64(0.99) * F64(d.courant_length(d.len_x_global, d.len_y_global, d.len_z_global,
                                                nx, ny, nz)) / F64(d.c))
    n0 = 0.25
    d.n0 = R(n0)
    L = float(d.len_x_global)
    d.name = f"dioctron_{nx}x{ny}x{nz}x{nppc}"
    d.plasma = dict(seed=seed, profile=1, ring_radius=0.3 * L, ring_width=0.05 * L,
                    column_charge=float(d.qsp) * n0, guide_bz=bz, pert_amp=amp, pert_mode=mode)
    return d.derive_params()


def column_mode_amplitude(deck, fields, mode=3):
    """Amplitude of the azimuthal mode `mode` of the radial electric field of
    an electron-column deck, about the column axis (the box centre in x, y):
        A = |sum over interior voxels of E_r(v) exp(-i mode theta(v))| / nz,
    with ex and ey each taken at its own Yee point (ex on the x edge at the
    (y-, z-) corner of voxel v, ey on the y edge at its (x-, z-) corner) and
    E_r = ex cos(theta) + ey sin(theta) there.  Returns A and dA/dfields
    (9 x Ng, FP64), the seed of its adjoint."""
    f = np.asarray(fields, np.float64).reshape(9, -1)
    nx, ny, nz = deck.nx, deck.ny, deck.nz
    dx, dy = float(deck.dx), float(deck.dy)
    X, Y = nx + 2, ny + 2
    ix = np.arange(1, nx + 1)
    iy = np.arange(1, ny + 1)
    seed = np.zeros_like(f)
    C = S = 0.0
    parts = []
    for comp, (ox, oy) in ((0, (0.5, 0.0)), (1, (0.0, 0.5))):
        xx = ((ix[None, :] - 1 + ox) * dx) - 0.5 * nx * dx
        yy = ((iy[:, None] - 1 + oy) * dy) - 0.5 * ny * dy
        th = np.arctan2(yy, xx)
        proj = np.cos(th) if comp == 0 else np.sin(th)
        gc = proj * np.cos(mode * th)  # (ny, nx)
        gs = -proj * np.sin(mode * th)
        vox = (ix[None, :] + X * iy[:, None])[None, :, :] + X * Y * np.arange(1, nz + 1)[:, None, None]
        e = f[comp][vox]  # (nz, ny, nx)
        C += float((e * gc[None]).sum())
        S += float((e * gs[None]).sum())
        parts.append((comp, vox, gc, gs))
    A = np.hypot(C, S) / nz
    for comp, vox, gc, gs in parts:
        seed[comp][vox] = np.broadcast_to((C * gc + S * gs)[None] / (A * nz * nz), vox.shape)
    return A, seed


def column_guide_field_tangent(deck, rel=1e-6):
    """The initial state's derivative with respect to the guide field B_z of
    an electron-column deck: cbz = B_z in every voxel (fields), and the
    E x B drift momenta of the particles, by central differences of the
    FP64 loader (positions and weights do not depend on B_z).  Returns
    (fields tangent (9, Ng), particle tangent (6, n))."""
    bz = deck.plasma["guide_bz"]
    vals = []
    for sgn in (1, -1):
        d = dioctron_3d(np.float64, nx=deck.nx, ny=deck.ny, nz=deck.nz, nppc=deck.nppc,
                        steps=deck.num_steps, seed=deck.plasma["seed"], bz=bz * (1 + sgn * rel),
                        amp=deck.plasma["pert_amp"], mode=deck.plasma["pert_mode"],
                        dx=float(deck.dx))
        vals.append(d.load_particles())
    t = np.zeros((6, vals[0]["cell"].size))
    for k, m in ((3, "ux"), (4, "uy"), (5, "uz")):
        t[k] = (vals[0][m] - vals[1][m]) / (2 * rel * bz)
    tf = np.zeros((9, deck.num_cells))
    tf[5] = 1.0
    return tf, t


CONFIGS = {
    "twostream1d": two_stream_1d,
    "weibel128": weibel_3d,
    "twostream64": two_stream_3d,
    "dioctron": dioctron_3d,
}
