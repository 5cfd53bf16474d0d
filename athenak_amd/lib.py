"""Python handle on libdustk.so (the gfx950 drag-coupling library) through its C-ABI.

torch supplies device memory and the stream; every computation runs in the library's HIP
kernels.  There is no fallback: without the built library or a HIP device this module raises.
"""

import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdustk.so")
_LIB = None

TASKS = ("first_two_imp_rk", "explicit_push", "update_particle_gids", "deposit_drag",
         "fold_deposit_qp", "gas_implicit_solve", "fill_ustar", "gather_kick_pmbr",
         "fold_pmbr", "apply_pmbr", "add_drag_history_gas", "fill_gas_ghosts")

EXPORTS = (["dk_create", "dk_destroy", "dk_bind_particles", "dk_bind_gas",
            "dk_export_particles", "dk_npart", "dk_stage_run", "dk_set_graphs",
            "dk_gas_hook", "dk_sort_particles", "dk_new_time_step", "dk_diag_sums",
            "dk_error", "dk_field", "dk_strerror", "dk_setup_problem", "dk_prof_enable",
            "dk_prof_reset", "dk_prof_read", "dk_autotune", "dk_tune_config",
            "dk_set_solver", "dk_solver_stats",
            "dk_set_tune_config", "dk_graph_launches", "dk_stiffness", "dk_keep_registers",
            "dk_stopping_times", "dk_remove_outflow_particles", "dk_x3_removed",
            "dk_sort_state", "dk_sort_particles_noalloc", "dk_write_pbin",
            "dk_particle_history", "dk_write_phst"] +
           ["dk_" + t for t in TASKS])


class DustError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise DustError("libdustk.so is not built (python -c 'import __graft_entry__ as g; "
                            "g.build()')")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.dk_create.argtypes = [C.POINTER(abi.dk_mesh), C.POINTER(abi.dk_block),
                                C.POINTER(abi.dk_physics), C.c_int32, P, C.POINTER(P)]
        L.dk_destroy.argtypes = [P]
        L.dk_bind_particles.argtypes = [P, P, P, C.c_int64, C.c_int64]
        L.dk_bind_gas.argtypes = [P, P, P]
        L.dk_export_particles.argtypes = [P]
        L.dk_npart.restype = C.c_int64
        L.dk_npart.argtypes = [P]
        st = C.POINTER(abi.dk_stage)
        for t in TASKS:
            getattr(L, "dk_" + t).argtypes = [P, st]
        L.dk_stage_run.argtypes = [P, st, C.POINTER(abi.dk_gas_src)]
        L.dk_gas_hook.argtypes = [P, st, C.POINTER(abi.dk_gas_src)]
        L.dk_set_graphs.argtypes = [P, C.c_int32]
        L.dk_stiffness.argtypes = [P, C.POINTER(C.c_double)]
        L.dk_stopping_times.argtypes = [P, C.POINTER(C.c_double)]
        L.dk_remove_outflow_particles.argtypes = [P, C.POINTER(C.c_double)]
        L.dk_x3_removed.argtypes = [P, C.POINTER(C.c_double)]
        L.dk_sort_state.argtypes = [P, C.c_int32, P, C.c_int64]
        L.dk_graph_launches.restype = C.c_int64
        L.dk_graph_launches.argtypes = [P]
        L.dk_autotune.argtypes = [P, C.c_int32]
        L.dk_set_solver.argtypes = [P, C.c_double, C.c_double, C.c_int32, C.c_double,
                                    C.c_double, C.c_double, C.c_double]
        L.dk_solver_stats.argtypes = [P, C.POINTER(C.c_double)]
        L.dk_tune_config.argtypes = [P] + [C.POINTER(C.c_int32)] * 3
        L.dk_set_tune_config.argtypes = [P, C.c_int32, C.c_int32]
        L.dk_keep_registers.argtypes = [P, C.c_int32]
        L.dk_sort_particles.argtypes = [P]
        L.dk_sort_particles_noalloc.argtypes = [P]
        L.dk_new_time_step.argtypes = [P, C.c_double, C.POINTER(C.c_double)]
        L.dk_diag_sums.argtypes = [P, C.POINTER(C.c_double)]
        L.dk_error.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.dk_field.restype = P
        L.dk_field.argtypes = [P, C.c_int32, C.POINTER(C.c_int32)]
        L.dk_strerror.restype = C.c_char_p
        L.dk_strerror.argtypes = [C.c_int]
        L.dk_setup_problem.argtypes = [P, C.POINTER(abi.dk_setup)]
        L.dk_prof_enable.argtypes = [P, C.c_int32]
        L.dk_prof_reset.argtypes = [P]
        L.dk_prof_read.argtypes = [P, C.c_int32, C.c_char_p, C.c_int32,
                                   C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        info = C.POINTER(abi.dk_output_info)
        L.dk_write_pbin.argtypes = [P, C.c_char_p, info]
        L.dk_particle_history.argtypes = [P, C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.dk_write_phst.argtypes = [P, C.c_char_p, info, C.POINTER(C.c_double), C.c_int32]
        _LIB = L
    return _LIB


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise DustError("no HIP device: the drag library runs on the GPU only")
    return torch


class DustContext:
    """A library context with its particle SoA and gas registers in device memory (torch
    tensors owned here, as the reference's Particles and Hydro own theirs)."""

    def __init__(self, mesh, blocks, phys, capacity, form=abi.DK_FORM_FP64, stream=None):
        torch = _torch()
        L = lib()
        self.torch = torch
        self.mesh, self.blocks, self.phys, self.form = mesh, blocks, phys, form
        self.cap = int(capacity)
        dev = torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self.rdata = torch.zeros((abi.DK_NRDATA, self.cap), dtype=torch.float64, device=dev)
        self.idata = torch.zeros((abi.DK_NIDATA, self.cap), dtype=torch.int32, device=dev)
        shape = (mesh.nmb, abi.DK_NGAS, mesh.nx3 + 2 * mesh.ng, mesh.nx2 + 2 * mesh.ng,
                 mesh.nx1 + 2 * mesh.ng)
        self.u0 = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.u1 = torch.zeros(shape, dtype=torch.float64, device=dev)
        # (torch filled the tensors on its current stream; the library works on self.stream)
        torch.cuda.synchronize()
        h = C.c_void_p()
        e = L.dk_create(C.byref(mesh), blocks, C.byref(phys), int(form),
                        C.c_void_p(self.stream.cuda_stream), C.byref(h))
        self._check(e, "dk_create")
        self.h = h
        self._check(L.dk_bind_gas(h, C.c_void_p(self.u0.data_ptr()),
                                  C.c_void_p(self.u1.data_ptr())), "dk_bind_gas")
        self.time = 0.0

    def close(self):
        if getattr(self, "h", None):
            lib().dk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _check(e, what):
        if e:
            raise DustError("%s: %s" % (what, lib().dk_strerror(e).decode()))

    def check_device_error(self):
        code, part = C.c_int32(0), C.c_int64(-1)
        self._check(lib().dk_error(self.h, C.byref(code), C.byref(part)), "dk_error")
        if code.value:
            raise DustError("device error: %s (particle %d)"
                            % (lib().dk_strerror(code.value).decode(), part.value))

    # ---- particles ----
    def bind_particles(self, npart):
        self.npart = int(npart)
        self._check(lib().dk_bind_particles(self.h, C.c_void_p(self.rdata.data_ptr()),
                                            C.c_void_p(self.idata.data_ptr()), self.cap,
                                            self.npart), "dk_bind_particles")

    def load(self, rdata, idata, u0, u1=None):
        """Copies host arrays (numpy, the oracle's layout) in and binds them."""
        torch = self.torch
        n = rdata.shape[1]
        self.rdata.zero_()
        self.idata.zero_()
        self.rdata[:, :n] = torch.from_numpy(rdata).to(self.rdata.device)
        self.idata[:, :n] = torch.from_numpy(idata).to(self.idata.device)
        self.u0.copy_(torch.from_numpy(u0))
        self.u1.copy_(torch.from_numpy(u1 if u1 is not None else u0))
        torch.cuda.synchronize()
        self.bind_particles(n)

    def setup(self, su):
        self._check(lib().dk_bind_particles(self.h, C.c_void_p(self.rdata.data_ptr()),
                                            C.c_void_p(self.idata.data_ptr()), self.cap, 0),
                    "dk_bind_particles")
        self._check(lib().dk_setup_problem(self.h, C.byref(su)), "dk_setup_problem")
        self.npart = lib().dk_npart(self.h)

    def export(self):
        """Writes the particles into rdata/idata; returns once they are there (the library's
        stream synchronised, so that torch work on any stream may read them)."""
        self._check(lib().dk_export_particles(self.h), "dk_export_particles")
        self.stream.synchronize()

    # ---- stepping ----
    def task(self, name, st):
        self._check(getattr(lib(), "dk_" + name)(self.h, C.byref(st)), name)

    def stage(self, st, src=None):
        self._check(lib().dk_stage_run(self.h, C.byref(st),
                                       C.byref(src) if src is not None else None),
                    "stage %d" % st.stage)

    def stages(self, dt):
        return (abi.imex2p_stages(self.time, dt) if self.phys.coupling == abi.DK_IMEX
                else abi.rk2_stages(self.time, dt))

    def step(self, dt, src=None, check=True):
        """One step's stages.  check: read the device error word afterwards (a host
        synchronisation; a timed loop passes False and checks once after it), so that an error
        recorded by a kernel (a cell beyond the sort's capacity, a particle outside the deposit
        halo) stops the run at the step where it happened."""
        for st in self.stages(dt):
            self.stage(st, src)
        self.time += dt
        if check:
            self.check_device_error()

    def set_graphs(self, on):
        self._check(lib().dk_set_graphs(self.h, 1 if on else 0), "dk_set_graphs")

    def stiffness(self):
        """UpdateStiffness records: chi of the last pc2 cycle, its running maximum, the rate
        maximum."""
        import numpy as np
        out = (C.c_double * 3)()
        self._check(lib().dk_stiffness(self.h, out), "dk_stiffness")
        return np.array(out[:])

    def stopping_times(self):
        """RefreshStoppingTimeMaximum: (largest, smallest) stopping time, validated."""
        out = (C.c_double * 2)()
        self._check(lib().dk_stopping_times(self.h, out), "dk_stopping_times")
        return out[0], out[1]

    def remove_outflow_particles(self):
        """RemoveOutflowParticles: (mass, count) removed now, then the run totals; the
        particle count drops by the count."""
        import numpy as np
        out = (C.c_double * 4)()
        self._check(lib().dk_remove_outflow_particles(self.h, out),
                    "dk_remove_outflow_particles")
        self.npart = lib().dk_npart(self.h)
        return np.array(out[:])

    def sort_state(self):
        """The FP32 form's last sort: (keys in the order before it, permutation new -> old
        slot, the first slot of every key), int32 numpy arrays."""
        import numpy as np
        m = self.mesh
        nk = m.nmb * m.nx1 * m.nx2 * m.nx3
        out = []
        for which, n in ((0, self.npart), (1, self.npart), (2, nk + 1)):
            a = np.zeros(max(n, 1), dtype=np.int32)
            self._check(lib().dk_sort_state(self.h, which, C.c_void_p(a.ctypes.data), n),
                        "dk_sort_state")
            out.append(a[:n])
        return tuple(out)

    def x3_removed(self):
        """The outflow ledger since the last removal: (mass, count)."""
        import numpy as np
        out = (C.c_double * 2)()
        self._check(lib().dk_x3_removed(self.h, out), "dk_x3_removed")
        return np.array(out[:])

    def new_time_step_pc2(self, dt_cfl, stiff_thresh=0.0):
        """NewTimeStep (dust_newdt.cpp:108-149) with the pc2 stiffness limit: min(dt_cfl min
        dx/|v|, stiff_thresh/stiff_rate_max) when stiff_thresh > 0 and a rate was recorded."""
        dt = self.new_time_step(dt_cfl)
        rate = self.stiffness()[2]
        if stiff_thresh > 0.0 and rate > 0.0 and self.phys.coupling == abi.DK_PC2:
            dt = min(dt, stiff_thresh / rate)
        return dt

    # ---- science outputs ----
    @staticmethod
    def output_info(time, dt, cycle=0, file_number=0, taus=(0.1,), stopping_time_mode=0,
                    single=False, basename="dust", output_block="output2",
                    source_commit=None, input_dump=""):
        """A dk_output_info (the strings are kept alive on the returned object)."""
        info = abi.dk_output_info(time=time, dt=dt, cycle=cycle, file_number=file_number,
                                  single_precision=int(bool(single)),
                                  stopping_time_mode=stopping_time_mode, nspecies=len(taus))
        for k, t in enumerate(taus):
            info.taus[k] = t
        info._keep = [x.encode() if isinstance(x, str) else x
                      for x in (basename, output_block, source_commit, input_dump)]
        info.basename, info.output_block, info.source_commit, info.input_dump = info._keep
        return info

    def write_pbin(self, path, info):
        """ParticleBinaryOutput: the pbin file (format version 2) of the current particles."""
        self._check(lib().dk_write_pbin(self.h, os.fsencode(path), C.byref(info)),
                    "dk_write_pbin")

    def particle_history(self, time, stiff_thresh=0.0):
        """DustGasDrag::ParticleHistory: the 14 values of a phst row after time and dt."""
        row = (C.c_double * 14)()
        self._check(lib().dk_particle_history(self.h, time, stiff_thresh, row),
                    "dk_particle_history")
        return list(row)

    def write_phst(self, path, info, row, header=False):
        r = (C.c_double * 14)(*row)
        self._check(lib().dk_write_phst(self.h, os.fsencode(path), C.byref(info), r,
                                        int(bool(header))), "dk_write_phst")

    def graphs_captured(self):
        """Stage graphs launched so far (0 on the null stream, where stages run eagerly)."""
        return lib().dk_graph_launches(self.h)

    def autotune(self, on=True):
        """Tune the FP32 form's launch parameters online over the next stages (or take them
        from the DK_TUNE_CACHE file); results are unchanged by any choice."""
        self._check(lib().dk_autotune(self.h, 1 if on else 0), "dk_autotune")

    def tune_config(self):
        """(state 0 fixed / 1 tuning / 2 tuned, xcd dealing, key-pass workgroup size)."""
        v = [C.c_int32(0) for _ in range(3)]
        self._check(lib().dk_tune_config(self.h, *[C.byref(x) for x in v]), "dk_tune_config")
        return tuple(x.value for x in v)

    def keep_registers(self, on):
        """Keep the stage registers x1, v1, R through every stage, so that an export after a step
        carries the reference's values (the FP32 form otherwise writes x1 = x, v1 = v, R = 0)."""
        self._check(lib().dk_keep_registers(self.h, 1 if on else 0), "dk_keep_registers")

    def set_tune_config(self, xcd, key_threads):
        self._check(lib().dk_set_tune_config(self.h, int(xcd), int(key_threads)),
                    "dk_set_tune_config")

    def set_solver(self, rtol=1e-11, atol=1e-14, iter_max=200, order_c=0.25, rtol_max=1e-3,
                   tref=1.0, state_floor=1.0):
        """The strict drag solver's parameters (the reference's <dust> defaults)."""
        self._check(lib().dk_set_solver(self.h, rtol, atol, iter_max, order_c, rtol_max, tref,
                                        state_floor), "dk_set_solver")

    def solver_stats(self):
        """The last implicit solve: iterations, fast accept, true residual, error bound,
        acceptance target, state scale, max Q/rho (-1 where not formed)."""
        out = (C.c_double * 7)()
        self._check(lib().dk_solver_stats(self.h, out), "dk_solver_stats")
        return list(out[:])

    def sort(self):
        self._check(lib().dk_sort_particles(self.h), "dk_sort_particles")

    def new_time_step(self, dt_cfl):
        out = C.c_double(0.0)
        self._check(lib().dk_new_time_step(self.h, dt_cfl, C.byref(out)), "dk_new_time_step")
        return out.value

    def diag(self):
        import numpy as np
        out = (C.c_double * 9)()
        self._check(lib().dk_diag_sums(self.h, out), "dk_diag_sums")
        return np.array(out[:])

    def prof(self, on=True):
        self._check(lib().dk_prof_enable(self.h, 1 if on else 0), "dk_prof_enable")

    def prof_reset(self):
        self._check(lib().dk_prof_reset(self.h), "dk_prof_reset")

    def prof_read(self):
        """{label: (total ms, launches)} of the kernels timed since the last reset."""
        out = {}
        k = 0
        buf = C.create_string_buffer(64)
        while True:
            ms, cnt = C.c_double(0), C.c_int64(0)
            if lib().dk_prof_read(self.h, k, buf, 64, C.byref(ms), C.byref(cnt)):
                break
            out[buf.value.decode()] = (ms.value, cnt.value)
            k += 1
        return out

    def field(self, which):
        """A host copy (numpy) of work field `which` (0 qdep, 1 ustar, 2 dmom, 3 ap,
        4 resid) in the gas layout."""
        import numpy as np
        nv = C.c_int32(0)
        p = lib().dk_field(self.h, int(which), C.byref(nv))
        if not p:
            return None
        m = self.mesh
        shape = (m.nmb, nv.value, m.nx3 + 2 * m.ng, m.nx2 + 2 * m.ng, m.nx1 + 2 * m.ng)
        host = np.empty(shape, dtype=np.float64)
        self.torch.cuda.synchronize()
        _copy_raw(host.ctypes.data, p, host.nbytes, kind=2)
        return host


def _copy_raw(dst, src, nbytes, kind=3):
    """hipMemcpy of raw pointers (kind 2: device to host, 3: device to device) through the
    HIP runtime that torch loads."""
    import ctypes
    hip = _hip()
    r = hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(nbytes),
                      kind)
    if r != 0:
        raise DustError("hipMemcpy failed (%d)" % r)


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        import ctypes
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _HIP = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise DustError("libamdhip64.so not found")
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int]
    return _HIP
