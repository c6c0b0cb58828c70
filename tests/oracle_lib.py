"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker for the HIP path and is
never imported by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

NCONSTS = 15
HIST_COLS = 6
DECK_NOUT = 33

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = ctypes.CDLL(ORACLE_SO)
        _declare(_lib)
    return _lib


def build_planted(define, outdir):
    """Builds the oracle with one ORA_PLANT_* switch on (a planted mistake)
    into outdir and returns it loaded.  Used only to show that the pinning
    checks fail on a wrong oracle."""
    out = os.path.join(outdir, f"liboracle_{define}.so")
    srcs = [os.path.join(ORACLE_DIR, f) for f in
            ("oracle_float.c", "oracle_double.c", "oracle_smooth.c")]
    subprocess.run(["gcc", "-O1", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-fPIC", "-shared", f"-D{define}", "-o", out] + srcs + ["-lm"], check=True)
    L = ctypes.CDLL(out)
    _declare(L)
    return L


def _declare(L):
    P = ctypes.c_void_p
    Lg = ctypes.c_long
    I = ctypes.c_int
    for sfx, R in (("_f", ctypes.c_float), ("_d", ctypes.c_double), ("_ds", ctypes.c_double)):
        g = lambda n: getattr(L, n + sfx)
        g("ora_load_interpolator").argtypes = [P, P, Lg, Lg, Lg]
        g("ora_clear_accumulator").argtypes = [P, Lg]
        g("ora_push").argtypes = [P] * 8 + [Lg, P, P] + [R] * 5 + [Lg, Lg, Lg, I]
        g("ora_push").restype = Lg
        g("ora_push_abs").argtypes = [P] * 8 + [Lg, P, P, P] + [R] * 5 + [Lg, Lg, Lg, I]
        g("ora_push_abs").restype = Lg
        g("ora_push_sum").argtypes = [P] * 8 + [Lg, P, P, P, P] + [R] * 5 + [Lg, Lg, Lg, I]
        g("ora_push_sum").restype = Lg
        g("ora_push_omp").argtypes = [P] * 8 + [Lg, P, P] + [R] * 5 + [Lg, Lg, Lg, I, I]
        g("ora_push_omp").restype = Lg
        g("ora_unload_accumulator").argtypes = [P, P, Lg, Lg, Lg, R, R, R, R]
        g("ora_uncenter").argtypes = [P] * 7 + [Lg, P, R, Lg, Lg, Lg]
        g("ora_fold_ghosts").argtypes = [P, P, P, Lg, Lg, Lg]
        g("ora_copy_ghosts").argtypes = [P, P, P, Lg, Lg, Lg]
        g("ora_advance_b").argtypes = [P, R, R, R, Lg, Lg, Lg]
        g("ora_advance_e").argtypes = [P, R, R, R, Lg, Lg, Lg, R]
        g("ora_energies").argtypes = [P, Lg, Lg, Lg, P, P, P, P]
        g("ora_es_advance_e").argtypes = [P, Lg, Lg, Lg, R]
        g("ora_es_energy").argtypes = [P, Lg, Lg, Lg, P, P]
        g("ora_kinetic_energy").argtypes = [P, P, P, P, Lg, R, R]
        g("ora_kinetic_energy").restype = ctypes.c_double
        g("ora_default_particle_init").argtypes = [P] * 8 + [Lg, Lg, Lg, R, Lg, R, R]
        g("ora_run").argtypes = [P] * 10 + [P, Lg, Lg, Lg, Lg, P, Lg, I, P]
        g("ora_run").restype = Lg
        g("ora_deck_2stream_em").argtypes = [P]
        g("ora_step_adjoint").argtypes = [P] * 9 + [Lg, Lg, Lg, Lg, P, P, P, P]
        g("ora_stage_adjoint").argtypes = [I, P, P, P, Lg, Lg, Lg] + [ctypes.c_double] * 6


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def sfx(dtype):
    return "_f" if np.dtype(dtype) == np.float32 else "_d"


class OracleState:
    """Host state of one simulation in the layouts the oracle expects."""

    def __init__(self, nx, ny, nz, np_, dtype):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dtype = np.dtype(dtype)
        self.ng = (nx + 2) * (ny + 2) * (nz + 2)
        self.fields = np.zeros((9, self.ng), self.dtype)
        self.interp = np.zeros((18, self.ng), self.dtype)
        self.acc = np.zeros((self.ng, 3, 4), self.dtype)
        self.p = {k: np.zeros(np_, self.dtype) for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
        self.cell = np.zeros(np_, np.int32)
        self.np = np_

    def copy(self):
        o = OracleState.__new__(OracleState)
        o.__dict__.update(self.__dict__)
        o.fields = self.fields.copy()
        o.interp = self.interp.copy()
        o.acc = self.acc.copy()
        o.p = {k: v.copy() for k, v in self.p.items()}
        o.cell = self.cell.copy()
        return o

    def _pp(self):
        return [_p(self.p[k]) for k in ("x", "y", "z", "ux", "uy", "uz", "w")] + [_p(self.cell)]

    def fn(self, name):
        # library: the oracle, or a planted variant (build_planted)
        L = getattr(self, "library", None) or lib()
        # smooth: the FP64 oracle with a double sqrt in place of the
        # reference's sqrtf (oracle/oracle_smooth.c), for finite differences
        if getattr(self, "smooth", False):
            assert self.dtype == np.float64
            return getattr(L, name + "_ds")
        return getattr(L, name + sfx(self.dtype))

    def load_interpolator(self):
        self.fn("ora_load_interpolator")(_p(self.fields), _p(self.interp), self.nx, self.ny, self.nz)

    def clear_accumulator(self):
        self.fn("ora_clear_accumulator")(_p(self.acc), self.ng)

    def push(self, qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp, boundary=1, nthreads=1, abs_acc=None,
             sum_acc=None):
        if sum_acc is not None:
            for a in (abs_acc, sum_acc):
                assert a is None or (a.dtype == np.float64 and a.size == 12 * self.ng)
            return self.fn("ora_push_sum")(*self._pp(), self.np, _p(self.interp), _p(self.acc),
                                           None if abs_acc is None else _p(abs_acc),
                                           _p(sum_acc), qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp,
                                           self.nx, self.ny, self.nz, boundary)
        if abs_acc is not None:
            assert abs_acc.dtype == np.float64 and abs_acc.size == 12 * self.ng
            return self.fn("ora_push_abs")(*self._pp(), self.np, _p(self.interp), _p(self.acc),
                                           _p(abs_acc), qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp,
                                           self.nx, self.ny, self.nz, boundary)
        if nthreads > 1:
            return self.fn("ora_push_omp")(*self._pp(), self.np, _p(self.interp), _p(self.acc),
                                           qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp,
                                           self.nx, self.ny, self.nz, boundary, nthreads)
        return self.fn("ora_push")(*self._pp(), self.np, _p(self.interp), _p(self.acc),
                                   qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp,
                                   self.nx, self.ny, self.nz, boundary)

    def uncenter(self, qdt_2mc):
        self.fn("ora_uncenter")(*self._pp()[:6], _p(self.cell), self.np, _p(self.interp),
                                qdt_2mc, self.nx, self.ny, self.nz)

    def unload(self, dx, dy, dz, dt):
        self.fn("ora_unload_accumulator")(_p(self.fields), _p(self.acc), self.nx, self.ny, self.nz,
                                          dx, dy, dz, dt)

    def advance_b(self, px, py, pz):
        self.fn("ora_advance_b")(_p(self.fields), px, py, pz, self.nx, self.ny, self.nz)

    def advance_e(self, px, py, pz, dt_eps0):
        self.fn("ora_advance_e")(_p(self.fields), px, py, pz, self.nx, self.ny, self.nz, dt_eps0)

    def fold_ghosts(self):
        f = self.fields
        self.fn("ora_fold_ghosts")(_p(f[6]), _p(f[7]), _p(f[8]), self.nx, self.ny, self.nz)

    def copy_ghosts(self, first):
        f = self.fields
        self.fn("ora_copy_ghosts")(_p(f[first]), _p(f[first + 1]), _p(f[first + 2]),
                                   self.nx, self.ny, self.nz)

    def es_advance_e(self, dt_eps0):
        self.fn("ora_es_advance_e")(_p(self.fields), self.nx, self.ny, self.nz, dt_eps0)

    def es_energy(self):
        e64 = ctypes.c_double()
        er = (ctypes.c_float if self.dtype == np.float32 else ctypes.c_double)()
        self.fn("ora_es_energy")(_p(self.fields), self.nx, self.ny, self.nz, ctypes.byref(e64),
                                 ctypes.byref(er))
        return e64.value, er.value

    def energies(self):
        e64 = ctypes.c_double()
        b64 = ctypes.c_double()
        R = ctypes.c_float if self.dtype == np.float32 else ctypes.c_double
        er, br = R(), R()
        self.fn("ora_energies")(_p(self.fields), self.nx, self.ny, self.nz,
                                ctypes.byref(e64), ctypes.byref(b64), ctypes.byref(er), ctypes.byref(br))
        return e64.value, b64.value, er.value, br.value

    def kinetic_energy(self, me=1.0, c=1.0):
        return self.fn("ora_kinetic_energy")(_p(self.p["ux"]), _p(self.p["uy"]), _p(self.p["uz"]),
                                             _p(self.p["w"]), self.np, me, c)

    def step_adjoint(self, consts, lf, lp_out):
        """Adjoint of one step (ora_step_adjoint) around this state, which is
        left unchanged: lf (9 x Ng doubles), the adjoint of the fields after
        the step, becomes that before it; lp_out (6 x np doubles) is the
        adjoint of the particles after the step; returns that before it."""
        k = np.asarray(consts, self.dtype)
        assert k.size == NCONSTS and lf.dtype == np.float64 and lf.size == 9 * self.ng
        lp_out = np.ascontiguousarray(lp_out, np.float64)
        assert lp_out.size == 6 * self.np
        lp_in = np.zeros((6, self.np))
        self.fn("ora_step_adjoint")(_p(self.fields), *self._pp(), self.np, self.nx, self.ny,
                                    self.nz, _p(k), _p(lf), _p(lp_out), _p(lp_in))
        return lp_in

    def run(self, consts, nsteps, nthreads=1, hist=True):
        k = np.asarray(consts, self.dtype)
        assert k.size == NCONSTS
        h = np.zeros((nsteps, HIST_COLS)) if hist else None
        moved = self.fn("ora_run")(_p(self.fields), _p(self.interp), _p(self.acc), *self._pp(),
                                   self.np, self.nx, self.ny, self.nz, _p(k), nsteps, nthreads,
                                   _p(h) if hist else None)
        return h, moved


def deck_2stream_em(dtype):
    """Constants of the 2stream-em regression deck, computed by the oracle."""
    out = np.zeros(DECK_NOUT)
    getattr(lib(), "ora_deck_2stream_em" + sfx(dtype))(_p(out))
    names = ["nx", "ny", "nz", "nppc", "num_steps", "np", "len_x", "len_y", "len_z",
             "dx", "dy", "dz", "dt", "n0", "v0", "Npe", "we", "dxp"]
    d = {n: out[i] for i, n in enumerate(names)}
    for n in ("nx", "ny", "nz", "nppc", "num_steps", "np"):
        d[n] = int(d[n])
    d["consts"] = out[len(names):]
    return d


def default_particle_init(st, dxp, nppc, we, v0):
    st.fn("ora_default_particle_init")(*st._pp(), st.np, st.nx, st.ny, dxp, nppc, we, v0)


def setup_2stream_em(dtype, library=None):
    """Initial state of the 2stream-em deck: zero fields, default particles."""
    d = deck_2stream_em(dtype)
    st = OracleState(d["nx"], d["ny"], d["nz"], d["np"], dtype)
    st.library = library
    default_particle_init(st, d["dxp"], d["nppc"], d["we"], d["v0"])
    return d, st
