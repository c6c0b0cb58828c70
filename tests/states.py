"""Seeded random simulation states shared by the parity tests.

A state is a small periodic grid with random fields on every voxel (ghosts
included) and particles in random interior cells.  The same host arrays are
handed to the oracle and to the HIP path.
"""
import numpy as np

import oracle_lib


def random_state(nx, ny, nz, ppc, dtype, seed, umax=0.6, cdt=(0.45, 0.4, 0.35)):
    rng = np.random.default_rng(seed)
    nc = nx * ny * nz
    np_ = nc * ppc
    st = oracle_lib.OracleState(nx, ny, nz, np_, dtype)
    st.fields[:] = rng.uniform(-1, 1, st.fields.shape).astype(dtype)
    c = np.repeat(np.arange(nc), ppc)
    rng.shuffle(c)
    ix = 1 + c % nx
    iy = 1 + (c // nx) % ny
    iz = 1 + c // (nx * ny)
    st.cell[:] = ix + (nx + 2) * (iy + (ny + 2) * iz)
    for k in ("x", "y", "z"):
        st.p[k][:] = rng.uniform(-1, 1, np_).astype(dtype)
    for k in ("ux", "uy", "uz"):
        st.p[k][:] = rng.uniform(-umax, umax, np_).astype(dtype)
    st.p["w"][:] = rng.uniform(0.5, 1.5, np_).astype(dtype)
    k = dict(qdt_2mc=dtype(-0.05), cdt_dx=dtype(cdt[0]), cdt_dy=dtype(cdt[1]),
             cdt_dz=dtype(cdt[2]), qsp=dtype(-1.0))
    return st, k


def step_consts(dtype, cdt=(0.45, 0.4, 0.35), dt=0.1):
    """Constants of a full step consistent with the cdt of random_state."""
    R = dtype
    dt = R(dt)
    dx, dy, dz = (R(dt / R(c)) for c in cdt)
    c = R(1)
    k = dict(qdt_2mc=R(R(-1) * dt / R(2)), cdt_dx=R(c * dt / dx), cdt_dy=R(c * dt / dy),
             cdt_dz=R(c * dt / dz), qsp=R(-1), dx=dx, dy=dy, dz=dz, dt=dt,
             px=R(c * dt / dx), py=R(c * dt / dy), pz=R(c * dt / dz), dt_eps0=dt,
             me=R(1), c=R(1))
    return k


def consts_vector(k):
    order = ["qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp", "dx", "dy", "dz", "dt",
             "px", "py", "pz", "dt_eps0", "me", "c"]
    return [k[n] for n in order]


def interior_mask(nx, ny, nz):
    m = np.zeros((nz + 2, ny + 2, nx + 2), bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.ravel()


MEMBERS = ("x", "y", "z", "ux", "uy", "uz", "w")


def canonical(d):
    """Particles ordered by (cell, bit patterns of x, y, z, ux, uy, uz, w):
    the HIP path keeps particles binned by cell, so particle sets are
    compared in this canonical order."""
    ui = {4: np.uint32, 8: np.uint64}[np.asarray(d["x"]).dtype.itemsize]
    keys = [np.ascontiguousarray(d[m]).view(ui) for m in reversed(MEMBERS)] + [np.asarray(d["cell"])]
    order = np.lexsort(keys)
    out = {m: np.ascontiguousarray(np.asarray(d[m])[order]) for m in MEMBERS}
    out["cell"] = np.ascontiguousarray(np.asarray(d["cell"])[order])
    return out


def oracle_particles(st):
    d = dict(st.p)
    d["cell"] = st.cell
    return d
