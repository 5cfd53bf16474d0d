"""Python handle on the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the package smoke check and by the CPU
baseline of bench.py, always as the checker, never as the measured or shipped path."""

import ctypes as C
import os

import numpy as np

from athenak_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        P = C.c_void_p
        L.or_create.restype = P
        L.or_create.argtypes = [C.POINTER(abi.dk_mesh), C.POINTER(abi.dk_block),
                                C.POINTER(abi.dk_physics), C.c_int64]
        L.or_destroy.argtypes = [P]
        L.or_set_threads.argtypes = [P, C.c_int]
        for n in ("or_rdata", "or_u0", "or_u1"):
            getattr(L, n).restype = C.POINTER(C.c_double)
            getattr(L, n).argtypes = [P]
        L.or_idata.restype = C.POINTER(C.c_int32)
        L.or_idata.argtypes = [P]
        for n in ("or_capacity", "or_npart", "or_cells_per_block"):
            getattr(L, n).restype = C.c_int64
            getattr(L, n).argtypes = [P]
        L.or_set_npart.argtypes = [P, C.c_int64]
        L.or_field.restype = C.POINTER(C.c_double)
        L.or_field.argtypes = [P, C.c_int, C.POINTER(C.c_int)]
        L.or_setup.argtypes = [P, C.POINTER(abi.dk_setup)]
        st = C.POINTER(abi.dk_stage)
        for n in ("or_first_two_imp_rk", "or_explicit_push", "or_update_particle_gids",
                  "or_deposit_drag", "or_fold_deposit_qp", "or_gas_implicit_solve",
                  "or_fill_ustar", "or_gather_kick_pmbr", "or_fold_pmbr", "or_apply_pmbr",
                  "or_add_drag_history_gas", "or_fill_gas_ghosts"):
            getattr(L, n).argtypes = [P, st]
        L.or_gas_hook.argtypes = [P, st, C.POINTER(abi.dk_gas_src)]
        L.or_stage_run.argtypes = [P, st, C.POINTER(abi.dk_gas_src)]
        L.or_sort_particles.argtypes = [P]
        L.or_new_time_step.argtypes = [P, C.c_double, C.POINTER(C.c_double)]
        L.or_apply_operator.argtypes = [P, C.c_double, C.c_double, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        L.or_diag_sums.argtypes = [P, C.POINTER(C.c_double)]
        L.or_last_error.argtypes = [P, C.POINTER(C.c_int64)]
        D = C.POINTER(C.c_double)
        L.or_cloud_cells.argtypes = [P, C.c_int, D, C.POINTER(C.c_int64), D, D]
        L.or_cloud_gather.argtypes = [P, D, C.c_int, D, D]
        L.or_leaf_cells.argtypes = [P, D, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                    D, D, C.c_int]
        L.or_leaf_gather.argtypes = [P, D, D, C.c_int, D]
        L.or_leaf_at.argtypes = [P, D]
        L.or_set_solver.argtypes = [P, C.c_double, C.c_double, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_double]
        L.or_solver_stats.argtypes = [P, D]
        L.or_stiffness.argtypes = [P, D]
        L.or_remove_outflow_particles.argtypes = [P, D]
        L.or_x3_removed.argtypes = [P, D]
        _LIB = L
    return _LIB


class Oracle:
    """One oracle simulation: mesh + particles + gas, stepped with the reference's stage
    sequence."""

    def __init__(self, mesh, blocks, phys, capacity, threads=1):
        L = lib()
        self.mesh, self.blocks, self.phys = mesh, blocks, phys
        self.h = L.or_create(C.byref(mesh), blocks, C.byref(phys), int(capacity))
        if not self.h:
            raise ValueError("or_create rejected the mesh")
        L.or_set_threads(self.h, int(threads))
        self.cap = L.or_capacity(self.h)
        self.ncb = L.or_cells_per_block(self.h)
        self.time = 0.0

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def set_threads(self, n):
        lib().or_set_threads(self.h, int(n))

    # ---- views (numpy arrays aliasing oracle storage) ----
    @property
    def npart(self):
        return lib().or_npart(self.h)

    def rdata(self):
        p = lib().or_rdata(self.h)
        return np.ctypeslib.as_array(p, shape=(abi.DK_NRDATA, self.cap))

    def idata(self):
        p = lib().or_idata(self.h)
        return np.ctypeslib.as_array(p, shape=(abi.DK_NIDATA, self.cap))

    def _field(self, ptr, nvar):
        m = self.mesh
        shape = (m.nmb, nvar, m.nx3 + 2 * m.ng, m.nx2 + 2 * m.ng, m.nx1 + 2 * m.ng)
        return np.ctypeslib.as_array(ptr, shape=shape)

    def u0(self):
        return self._field(lib().or_u0(self.h), abi.DK_NGAS)

    def u1(self):
        return self._field(lib().or_u1(self.h), abi.DK_NGAS)

    def field(self, which):
        nv = C.c_int(0)
        p = lib().or_field(self.h, int(which), C.byref(nv))
        return self._field(p, nv.value)

    # ---- set-up and stepping ----
    def setup(self, su):
        e = lib().or_setup(self.h, C.byref(su))
        if e:
            raise RuntimeError("or_setup: %s" % abi.ERRORS.get(e, e))

    def check(self, e, what):
        if e:
            pp = C.c_int64(-1)
            code = lib().or_last_error(self.h, C.byref(pp))
            raise RuntimeError("%s: %s (particle %d)" % (what, abi.ERRORS.get(e, e), pp.value))

    def task(self, name, st, *extra):
        e = getattr(lib(), "or_" + name)(self.h, C.byref(st), *extra)
        self.check(e, name)

    def stage(self, st, src=None):
        e = lib().or_stage_run(self.h, C.byref(st), C.byref(src) if src is not None else None)
        self.check(e, "stage %d" % st.stage)

    def step(self, dt, src=None):
        stages = (abi.imex2p_stages(self.time, dt) if self.phys.coupling == abi.DK_IMEX
                  else abi.rk2_stages(self.time, dt))
        for st in stages:
            self.stage(st, src)
        self.time += dt

    def sort(self):
        lib().or_sort_particles(self.h)

    def diag(self):
        out = (C.c_double * 9)()
        lib().or_diag_sums(self.h, out)
        return np.array(out[:])

    def cloud_cells(self, m, x):
        """(padded cell indices, weights, first moments [n, 3]) under the cloud at x."""
        xs = (C.c_double * 3)(*x)
        cell = (C.c_int64 * 27)()
        w = (C.c_double * 27)()
        m1 = (C.c_double * 81)()
        n = lib().or_cloud_cells(self.h, m, xs, cell, w, m1)
        assert n >= 0, "cloud beyond the first ghost layer"
        return (np.array(cell[:n]), np.array(w[:n]), np.array(m1[:3 * n]).reshape(n, 3))

    def cloud_gather(self, vel, m, x):
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        xs = (C.c_double * 3)(*x)
        out = (C.c_double * 3)()
        assert lib().or_cloud_gather(self.h, vel.ctypes.data_as(C.POINTER(C.c_double)), m,
                                     xs, out) == 0
        return np.array(out[:])

    def leaf_cells(self, x, lev, cap=4096):
        """On a refined mesh: (blocks, padded cell indices, weights, first moments [n, 3]) of the
        leaf cells under the cloud of level lev at x."""
        xs = (C.c_double * 3)(*x)
        blk = (C.c_int32 * cap)()
        cell = (C.c_int64 * cap)()
        w = (C.c_double * cap)()
        m1 = (C.c_double * (3 * cap))()
        n = lib().or_leaf_cells(self.h, xs, int(lev), blk, cell, w, m1, cap)
        assert n >= 0, "cloud support not covered"
        return (np.array(blk[:n]), np.array(cell[:n]), np.array(w[:n]),
                np.array(m1[:3 * n]).reshape(n, 3))

    def leaf_gather(self, vel, x, lev):
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        xs = (C.c_double * 3)(*x)
        out = (C.c_double * 3)()
        assert lib().or_leaf_gather(self.h, vel.ctypes.data_as(C.POINTER(C.c_double)), xs,
                                    int(lev), out) == 0
        return np.array(out[:])

    def leaf_at(self, x):
        return lib().or_leaf_at(self.h, (C.c_double * 3)(*x))

    def set_solver(self, rtol=1e-11, atol=1e-14, iter_max=200, order_c=0.25, rtol_max=1e-3,
                   tref=1.0, state_floor=1.0):
        """The strict solver's parameters (the reference's <dust> defaults)."""
        lib().or_set_solver(self.h, rtol, atol, iter_max, order_c, rtol_max, tref, state_floor)

    def solver_stats(self):
        """The last implicit solve: iterations, fast accept, true residual, error bound,
        acceptance target, state scale, max Q/rho (-1 where not formed)."""
        out = (C.c_double * 7)()
        lib().or_solver_stats(self.h, out)
        return np.array(out[:])

    def stiffness(self):
        """UpdateStiffness records: chi of the last pc2 cycle, its running maximum, the rate
        maximum."""
        out = (C.c_double * 3)()
        lib().or_stiffness(self.h, out)
        return np.array(out[:])

    def remove_outflow_particles(self):
        """RemoveOutflowParticles: (mass, count) removed now, (mass, count) run totals."""
        out = (C.c_double * 4)()
        lib().or_remove_outflow_particles(self.h, out)
        return np.array(out[:])

    def x3_removed(self):
        """The outflow ledger since the last removal: (mass, count)."""
        out = (C.c_double * 2)()
        lib().or_x3_removed(self.h, out)
        return np.array(out[:])

    def apply_operator(self, a_dt, x, sh=0.0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().or_apply_operator(self.h, a_dt, sh, x.ctypes.data_as(C.POINTER(C.c_double)),
                                y.ctypes.data_as(C.POINTER(C.c_double)))
        return y
