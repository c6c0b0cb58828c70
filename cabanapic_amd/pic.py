"""Host-side mirror of the reference's hot-path interface.

``Simulation`` owns one device context (particles, fields, interpolators and
accumulators resident in HBM) and exposes the reference's stage functions
under their reference names, with the same argument meaning:

=============================  ==============================================
reference                      here
=============================  ==============================================
load_interpolator_array        Simulation.load_interpolator_array()
  (src/interpolator.h:14-21)
clear_accumulator_array        Simulation.clear_accumulator_array()
  (src/accumulator.h:15-21)
push (src/push.h:7-23)         Simulation.push(qdt_2mc, cdt_dx, cdt_dy,
                               cdt_dz, qsp, boundary)
unload_accumulator_array       Simulation.unload_accumulator_array(dx, dy, dz,
  (src/accumulator.h:23-34)    dt)
Field_Solver::advance_b        Simulation.advance_b(px, py, pz)
  (src/fields.h:352-364)
Field_Solver::advance_e        Simulation.advance_e(px, py, pz, dt_eps0)
  (src/fields.h:365-378)
e_energy / b_energy            Simulation.e_energy(), .b_energy()
  (src/fields.h:556-615)
main loop body                 Simulation.step(params, nsteps)
  (example/example.cpp:216-279)
=============================  ==============================================

Errors raise ``CpicError``; the reference's ``exit(1)`` on a Reflect boundary
(src/fields.h:21-25) becomes CPIC_ERR_UNSUPPORTED.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import BOUNDARY_PERIODIC, CpicError, StepParams, check

GAUSS_BASELINE, GAUSS_CHANGE = 0, 1  # cpic_gauss_residual modes

PARTICLE_MEMBERS = ("x", "y", "z", "ux", "uy", "uz", "w")
FIELD_MEMBERS = ("ex", "ey", "ez", "cbx", "cby", "cbz", "jfx", "jfy", "jfz")
INTERP_MEMBERS = ("ex", "dexdy", "dexdz", "d2exdydz", "ey", "deydz", "deydx", "d2eydzdx",
                  "ez", "dezdx", "dezdy", "d2ezdxdy", "cbx", "dcbxdx", "cby", "dcbydy",
                  "cbz", "dcbzdz")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Simulation:
    """One simulation's device state on the current HIP device."""

    def __init__(self, nx, ny, nz, max_particles, dtype=np.float32, num_ghosts=1,
                 column_rows=0):
        self.lib = _capi.load()
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise ValueError("dtype must be float32 or float64")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.ncells = (self.nx + 2) * (self.ny + 2) * (self.nz + 2)
        cfg = _capi.Config(self.dtype.itemsize, num_ghosts, self.nx, self.ny, self.nz,
                           int(max_particles), int(column_rows))
        h = ctypes.c_void_p()
        check(self.lib.cpic_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._owned = True

    @classmethod
    def _view(cls, handle, nx, ny, nz, dtype):
        """A Simulation over a context owned by another object."""
        obj = cls.__new__(cls)
        obj.lib = _capi.load()
        obj.dtype = np.dtype(dtype)
        obj.nx, obj.ny, obj.nz = int(nx), int(ny), int(nz)
        obj.ncells = (obj.nx + 2) * (obj.ny + 2) * (obj.nz + 2)
        obj._h = handle
        obj._owned = False
        return obj

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                self.lib.cpic_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        check(self.lib.cpic_synchronize(self._h))

    # -- data movement ----------------------------------------------------------
    def _arr(self, a, dtype=None):
        return np.ascontiguousarray(a, dtype=self.dtype if dtype is None else dtype)

    def upload_particles(self, x, y, z, ux, uy, uz, w, cell):
        arrs = [self._arr(a) for a in (x, y, z, ux, uy, uz, w)]
        c = self._arr(cell, np.int32)
        n = c.size
        if any(a.size != n for a in arrs):
            raise ValueError("particle arrays differ in length")
        check(self.lib.cpic_upload_particles(self._h, n, *[_ptr(a) for a in arrs], _ptr(c)))

    def num_particles(self):
        n = ctypes.c_int64()
        check(self.lib.cpic_num_particles(self._h, ctypes.byref(n)))
        return n.value

    def download_particles(self):
        n = self.num_particles()
        out = {k: np.empty(n, self.dtype) for k in PARTICLE_MEMBERS}
        out["cell"] = np.empty(n, np.int32)
        check(self.lib.cpic_download_particles(
            self._h, n, *[_ptr(out[k]) for k in PARTICLE_MEMBERS], _ptr(out["cell"])))
        return out

    def download_particle_tags(self):
        """Upload index of every particle, in download_particles order."""
        n = self.num_particles()
        t = np.empty(n, np.int64)
        check(self.lib.cpic_download_particle_tags(self._h, n, _ptr(t)))
        return t

    def column_rows(self):
        r = ctypes.c_int64()
        check(self.lib.cpic_column_rows(self._h, ctypes.byref(r)))
        return r.value

    def reserve_column_rows(self, rows):
        """Grows the particle columns to at least `rows` rows on the device."""
        check(self.lib.cpic_reserve_column_rows(self._h, int(rows)))

    def upload_fields(self, fields):
        f = self._arr(fields)
        if f.size != 9 * self.ncells:
            raise ValueError("fields must hold 9 x Ng values")
        check(self.lib.cpic_upload_fields(self._h, _ptr(f)))

    def download_fields(self):
        f = np.empty((9, self.ncells), self.dtype)
        check(self.lib.cpic_download_fields(self._h, _ptr(f)))
        return f

    def upload_interpolators(self, interp):
        a = self._arr(interp)
        if a.size != 18 * self.ncells:
            raise ValueError("interpolators must hold 18 x Ng values")
        check(self.lib.cpic_upload_interpolators(self._h, _ptr(a)))

    def download_interpolators(self):
        a = np.empty((18, self.ncells), self.dtype)
        check(self.lib.cpic_download_interpolators(self._h, _ptr(a)))
        return a

    def upload_accumulators(self, acc):
        a = self._arr(acc)
        if a.size != 12 * self.ncells:
            raise ValueError("accumulators must hold Ng x 3 x 4 values")
        check(self.lib.cpic_upload_accumulators(self._h, _ptr(a)))

    def download_accumulators(self):
        a = np.empty((self.ncells, 3, 4), self.dtype)
        check(self.lib.cpic_download_accumulators(self._h, _ptr(a)))
        return a

    # -- hot-path stages --------------------------------------------------------
    def load_interpolator_array(self):
        check(self.lib.cpic_load_interpolator_array(self._h))

    def clear_accumulator_array(self):
        check(self.lib.cpic_clear_accumulator_array(self._h))

    def push(self, qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp, boundary=BOUNDARY_PERIODIC):
        check(self.lib.cpic_push(self._h, float(qdt_2mc), float(cdt_dx), float(cdt_dy),
                                 float(cdt_dz), float(qsp), int(boundary)))

    def uncenter_particles(self, qdt_2mc):
        check(self.lib.cpic_uncenter_particles(self._h, float(qdt_2mc)))

    def unload_accumulator_array(self, dx, dy, dz, dt):
        check(self.lib.cpic_unload_accumulator_array(self._h, float(dx), float(dy), float(dz),
                                                     float(dt)))

    def advance_b(self, px, py, pz):
        check(self.lib.cpic_advance_b(self._h, float(px), float(py), float(pz)))

    def advance_e(self, px, py, pz, dt_eps0):
        check(self.lib.cpic_advance_e(self._h, float(px), float(py), float(pz), float(dt_eps0)))

    def set_field_solver(self, solver):
        """"em" (default) or "es_1d" (ES_Field_Solver_1D)."""
        code = {"em": _capi.FIELD_SOLVER_EM, "es_1d": _capi.FIELD_SOLVER_ES_1D}[solver]
        check(self.lib.cpic_set_field_solver(self._h, code))

    def set_compensated_deposit(self, enabled):
        """Compensated (TwoSum) summation of each cell's current sums inside
        the push passes; off by default."""
        check(self.lib.cpic_set_compensated_deposit(self._h, 1 if enabled else 0))

    def update_ghosts_j(self):
        check(self.lib.cpic_update_ghosts_j(self._h))

    def energies(self):
        e, b = ctypes.c_double(), ctypes.c_double()
        check(self.lib.cpic_energies(self._h, ctypes.byref(e), ctypes.byref(b)))
        return e.value, b.value

    def gauss_residual(self, dx, dy, dz, eps0, qsp, baseline, residual=False):
        """Gauss's law check in FP64 (cpic_gauss_residual): with baseline
        True stores the residual div E - rho / eps0 of every interior node
        and returns (max |r|, max |rho / eps0|), else returns (max |r -
        baseline|, max |rho / eps0|); with residual True also the residual
        array, shape (nz, ny, nx)."""
        out = np.zeros(2)
        r = np.empty((self.nz, self.ny, self.nx)) if residual else None
        check(self.lib.cpic_gauss_residual(
            self._h, float(dx), float(dy), float(dz), float(eps0), float(qsp),
            GAUSS_BASELINE if baseline else GAUSS_CHANGE, _ptr(out),
            _ptr(r) if residual else None))
        return (out[0], out[1], r) if residual else (out[0], out[1])

    def e_energy(self):
        return self.energies()[0]

    def b_energy(self):
        return self.energies()[1]

    def kinetic_energy(self, me=1.0, c=1.0):
        ke = ctypes.c_double()
        check(self.lib.cpic_kinetic_energy(self._h, float(me), float(c), ctypes.byref(ke)))
        return ke.value

    def sort_particles(self):
        check(self.lib.cpic_sort_particles(self._h))

    def step(self, params, nsteps, history=True):
        """Runs nsteps of the reference loop body; returns the (nsteps, 3)
        energy history [E, B, KE] (NaN where no dump was taken)."""
        if isinstance(params, dict):
            params = make_step_params(**params)
        hist = np.full((int(nsteps), 3), np.nan) if history else None
        check(self.lib.cpic_step(self._h, ctypes.byref(params), int(nsteps),
                                 _ptr(hist) if history else None))
        return hist

    def step_async(self, params, nsteps):
        """Queues nsteps of the loop body without waiting for them (no
        history; device errors surface at the next synchronous call)."""
        if isinstance(params, dict):
            params = make_step_params(**params)
        check(self.lib.cpic_step_async(self._h, ctypes.byref(params), int(nsteps)))

    def snapshot_begin(self, slot):
        """Queues a dump snapshot into slot 0 or 1 (see snapshot_wait)."""
        check(self.lib.cpic_snapshot_begin(self._h, int(slot)))

    def snapshot_wait(self, slot):
        """Waits for the snapshot of the slot; returns copies of its data:
        x (in-cell offset), ux and cell of every particle in upload order,
        ex_line (ex at voxels (1..nx, 1, 1)) and the last energies."""
        sn = _capi.Snapshot()
        check(self.lib.cpic_snapshot_wait(self._h, int(slot), ctypes.byref(sn)))
        n, nx = int(sn.n), int(sn.nx)
        rt = ctypes.c_float if sn.real_size == 4 else ctypes.c_double

        def arr(ptr, count, ct):
            if count == 0:
                return np.zeros(0, np.dtype(ct))
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), (count,)).copy()

        return {"x": arr(sn.x, n, rt), "ux": arr(sn.ux, n, rt),
                "cell": arr(sn.cell, n, ctypes.c_int32), "ex_line": arr(sn.ex_line, nx, rt),
                "e_energy": sn.e_energy, "b_energy": sn.b_energy,
                "kinetic_energy": sn.kinetic_energy}

    # -- diagnostics ------------------------------------------------------------
    def set_timing(self, enabled, stages=None):
        """Times every stage (enabled), none, or the named stages only."""
        if stages is None:
            check(self.lib.cpic_set_timing(self._h, 1 if enabled else 0))
        else:
            mask = sum(1 << _capi.STAGES.index(s) for s in stages) if enabled else 0
            check(self.lib.cpic_set_timing_stages(self._h, mask))

    def stage_times(self):
        counts = np.zeros(_capi.NUM_STAGES, np.int64)
        ms = np.zeros(_capi.NUM_STAGES, np.float64)
        check(self.lib.cpic_stage_times(self._h, _ptr(counts), _ptr(ms)))
        return {name: (int(counts[i]), float(ms[i])) for i, name in enumerate(_capi.STAGES)}

    def last_push_movers(self):
        n = ctypes.c_int64()
        check(self.lib.cpic_last_push_movers(self._h, ctypes.byref(n)))
        return n.value

    def device_errors(self):
        f = ctypes.c_int32()
        check(self.lib.cpic_device_errors(self._h, ctypes.byref(f)))
        return f.value


class TangentSimulation:
    """The step and its forward-mode derivative with respect to one deck
    parameter (the cpic_tangent_* entry points).  `primal` and `tangent` are
    Simulation views of the two field contexts (fields, interpolators,
    accumulators, energies); the particles and their tangents live here."""

    def __init__(self, nx, ny, nz, max_particles, dtype=np.float32, num_ghosts=1):
        self.lib = _capi.load()
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise ValueError("dtype must be float32 or float64")
        cfg = _capi.Config(self.dtype.itemsize, num_ghosts, int(nx), int(ny), int(nz),
                           int(max_particles))
        h = ctypes.c_void_p()
        check(self.lib.cpic_tangent_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        p, t = ctypes.c_void_p(), ctypes.c_void_p()
        check(self.lib.cpic_tangent_contexts(h, ctypes.byref(p), ctypes.byref(t)))
        self.primal = Simulation._view(p, nx, ny, nz, self.dtype)
        self.tangent = Simulation._view(t, nx, ny, nz, np.float64)
        self._n = 0

    def close(self):
        if getattr(self, "_h", None):
            self.primal._h = self.tangent._h = None
            self.lib.cpic_tangent_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload_particles(self, x, y, z, ux, uy, uz, w, cell, tangent):
        """tangent: (6, n) derivatives of x, y, z, ux, uy, uz."""
        arrs = [np.ascontiguousarray(a, self.dtype) for a in (x, y, z, ux, uy, uz, w)]
        c = np.ascontiguousarray(cell, np.int32)
        n = c.size
        t = np.ascontiguousarray(tangent, np.float64)
        if any(a.size != n for a in arrs) or t.shape != (6, n):
            raise ValueError("particle arrays differ in length")
        check(self.lib.cpic_tangent_upload_particles(self._h, n, *[_ptr(a) for a in arrs],
                                                     _ptr(c), _ptr(t)))
        self._n = n

    def download_particles(self):
        n = self._n
        out = {k: np.empty(n, self.dtype) for k in PARTICLE_MEMBERS}
        out["cell"] = np.empty(n, np.int32)
        out["tangent"] = np.empty((6, n), np.float64)
        check(self.lib.cpic_tangent_download_particles(
            self._h, n, *[_ptr(out[k]) for k in PARTICLE_MEMBERS], _ptr(out["cell"]),
            _ptr(out["tangent"])))
        return out

    def push(self, qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp, boundary=BOUNDARY_PERIODIC):
        check(self.lib.cpic_tangent_push(self._h, float(qdt_2mc), float(cdt_dx),
                                         float(cdt_dy), float(cdt_dz), float(qsp),
                                         int(boundary)))

    def step(self, params, nsteps):
        """Runs nsteps; returns the (nsteps, 4) history [E, B, dE, dB]."""
        if isinstance(params, dict):
            params = make_step_params(**params)
        hist = np.full((int(nsteps), 4), np.nan)
        check(self.lib.cpic_tangent_step(self._h, ctypes.byref(params), int(nsteps),
                                         _ptr(hist)))
        return hist

    def last_push_movers(self):
        n = ctypes.c_int64()
        check(self.lib.cpic_tangent_last_push_movers(self._h, ctypes.byref(n)))
        return n.value

    def set_graph(self, enabled):
        """Replay the step as a HIP graph (default) or launch every stage."""
        check(self.lib.cpic_tangent_set_graph(self._h, 1 if enabled else 0))


class Adjoint:
    """Reverse-mode derivative of a run of steps on a Simulation (the
    cpic_forward / cpic_backward pair): forward() runs the steps keeping a
    checkpoint every `interval` steps; backward() takes the derivative of an
    objective with respect to the final state and returns it with respect to
    the initial state."""

    def __init__(self, sim):
        self.sim = sim
        self.lib = sim.lib
        h = ctypes.c_void_p()
        check(self.lib.cpic_adjoint_create(sim._h, ctypes.byref(h)))
        self._h = h
        self._n = 0

    def close(self):
        if getattr(self, "_h", None):
            self.lib.cpic_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def forward(self, params, nsteps, interval, history=True):
        """Runs nsteps steps; returns the (nsteps, 3) energy history [E, B,
        kinetic] as params' intervals ask (NaN elsewhere), or None."""
        if isinstance(params, dict):
            params = make_step_params(**params)
        hist = np.full((int(nsteps), 3), np.nan) if history else None
        check(self.lib.cpic_forward(self._h, ctypes.byref(params), int(nsteps), int(interval),
                                    _ptr(hist) if history else None))
        self._n = self.sim.num_particles()
        return hist

    def backward(self, seed_fields, seed_particles=None):
        """seed_fields: (9, Ng) derivative of the objective with respect to
        the final fields; seed_particles: (6, n) with respect to the final
        x, y, z, ux, uy, uz of each particle in upload order (or None).
        Returns (grad_fields (9, Ng), grad_particles (6, n)) with respect to
        the initial state."""
        sf = np.ascontiguousarray(seed_fields, np.float64)
        if sf.size != 9 * self.sim.ncells:
            raise ValueError("seed_fields must hold 9 x Ng values")
        sp = None
        if seed_particles is not None:
            sp = np.ascontiguousarray(seed_particles, np.float64)
            if sp.size != 6 * self._n:
                raise ValueError("seed_particles must hold 6 x n values")
        gf = np.empty((9, self.sim.ncells))
        gp = np.empty((6, self._n))
        check(self.lib.cpic_backward(self._h, _ptr(sf), _ptr(sp) if sp is not None else None,
                                     _ptr(gf), _ptr(gp)))
        return gf, gp

    def memory(self):
        """Device bytes: checkpoints (with the final state), one segment's
        snapshots, the adjoint arrays."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.cpic_adjoint_memory(self._h, ctypes.byref(a), ctypes.byref(b),
                                           ctypes.byref(c)))
        return {"checkpoints": a.value, "segment": b.value, "adjoint": c.value}


def make_step_params(qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp, dx, dy, dz, dt, px, py, pz,
                     dt_eps0, me=1.0, c=1.0, boundary=BOUNDARY_PERIODIC, energy_interval=1,
                     compact_interval=0, kinetic_interval=0):
    return StepParams(float(qdt_2mc), float(cdt_dx), float(cdt_dy), float(cdt_dz), float(qsp),
                      float(dx), float(dy), float(dz), float(dt), float(px), float(py),
                      float(pz), float(dt_eps0), float(me), float(c), int(boundary),
                      int(energy_interval), int(compact_interval), int(kinetic_interval))


__all__ = ["Simulation", "TangentSimulation", "make_step_params", "CpicError", "PARTICLE_MEMBERS", "FIELD_MEMBERS",
           "INTERP_MEMBERS"]
