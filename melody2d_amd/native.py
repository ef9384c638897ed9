"""ctypes binding of the C-ABI in include/melody2d.h.

Loads melody2d_amd/libmelody2d_hip.so (built in-tree by melody2d_amd.build).
There is no fallback: if the library is missing or cannot be loaded, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmelody2d_hip.so")

MEL_OK = 0
STATUS_NAMES = {0: "MEL_OK", 1: "MEL_E_ARG", 2: "MEL_E_IO", 3: "MEL_E_PARSE", 4: "MEL_E_OOM", 5: "MEL_E_HIP",
                6: "MEL_E_NAN", 7: "MEL_E_CAPACITY", 8: "MEL_E_UNSUPPORTED",
                9: "MEL_E_TIMEOUT"}
FP64, FP32_MIXED = 0, 1
EVAL_INTERNAL, EVAL_CONTACT, EVAL_EXTERNAL, EVAL_ALL = 1, 2, 4, 7

NODE_FIELDS = ["x_current", "y_current", "x_disp", "y_disp", "x_vel", "y_vel", "x_acc", "y_acc",
               "x_disp_param", "y_disp_param", "x_vel_param", "y_vel_param", "x_acc_param", "y_acc_param",
               "x_internal_force", "y_internal_force", "x_contact_force", "y_contact_force",
               "x_body_force", "y_body_force", "x_dirichlet_force", "y_dirichlet_force",
               "x_neumann_force", "y_neumann_force", "x_damping_force", "y_damping_force",
               "x_force", "y_force", "x_mass", "y_mass", "x_alid_force", "y_alid_force"]


class MelError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


class Info(ctypes.Structure):
    _fields_ = [("n_bodies", ctypes.c_int), ("n_nodes", ctypes.c_int), ("n_gauss", ctypes.c_int),
                ("n_borders", ctypes.c_int), ("n_border_nodes", ctypes.c_int), ("n_rigid", ctypes.c_int),
                ("n_contacts", ctypes.c_int), ("contact_capacity", ctypes.c_int),
                ("nnz_gauss", ctypes.c_longlong), ("nnz_d2d", ctypes.c_longlong),
                ("time", ctypes.c_double), ("dt", ctypes.c_double), ("iteration", ctypes.c_int),
                ("precision", ctypes.c_int)]


JOB_WAITING, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_RELEASED = range(5)


class JobInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("status", ctypes.c_int), ("steps", ctypes.c_longlong),
                ("steps_done", ctypes.c_longlong), ("arrival", ctypes.c_double), ("admitted", ctypes.c_double),
                ("finished", ctypes.c_double), ("device_bytes", ctypes.c_size_t), ("n_nodes", ctypes.c_int)]


# mel_stress_field order
STRESS_FIELDS = ["sigma_I", "sigma_II", "sigma_III", "jacobian", "Sxx", "Syy", "Sxy", "Szz", "sigma_xx", "sigma_yy",
                 "sigma_xy", "sigma_zz", "sigma_tresca", "sigma_vm", "sigma_sph", "Exx", "Eyy", "Exy", "norm_E"]


class SolverInfo(ctypes.Structure):
    """mel_solver_info_t"""
    _fields_ = [("scheme", ctypes.c_int), ("dt", ctypes.c_double), ("neval", ctypes.c_int),
                ("attempts", ctypes.c_longlong), ("max_error", ctypes.c_double), ("mean_error", ctypes.c_double),
                ("total_mass", ctypes.c_double), ("max_mass", ctypes.c_double)]


class Options(ctypes.Structure):
    """mel_options: load-time launch configuration (-1 = library default)."""
    _fields_ = [("block_nodes", ctypes.c_int), ("fused_step", ctypes.c_int),
                ("graph_replay", ctypes.c_int), ("wide_window", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def make_options(block_nodes=-1, fused_step=-1, graph_replay=-1, wide_window=-1):
    return Options(block_nodes, fused_step, graph_replay, wide_window)


_lib = None


def library():
    """The loaded HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("HIP library %s is not built; run melody2d_amd.build.build_all()" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        sigs = {
            "mel_load": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(vp)], ctypes.c_int),
            "mel_load_binary": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P(vp)], ctypes.c_int),
            "mel_load_ex": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(Options), P(vp)],
                            ctypes.c_int),
            "mel_load_binary_ex": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P(Options), P(vp)], ctypes.c_int),
            "mel_default_options": ([P(Options)], None),
            "mel_get_options": ([vp, P(Options)], ctypes.c_int),
            "mel_autotune": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(Options),
                              P(ctypes.c_double), P(ctypes.c_int)], ctypes.c_int),
            "mel_update_proximity": ([vp], ctypes.c_int),
            "mel_euler_steps": ([vp, ctypes.c_int], ctypes.c_int),
            "mel_synchronize": ([vp], ctypes.c_int),
            "mel_solver_steps": ([vp, ctypes.c_int], ctypes.c_int),
            "mel_get_solver_info": ([vp, P(SolverInfo)], ctypes.c_int),
            "mel_get_body_error": ([vp, ctypes.c_int, P(ctypes.c_double)], ctypes.c_int),
            "mel_get_mass_scaling": ([vp, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double), ctypes.c_size_t],
                                     ctypes.c_int),
            "mel_get_node_field": ([vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_double), ctypes.c_size_t], ctypes.c_int),
            "mel_set_node_field": ([vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_double), ctypes.c_size_t], ctypes.c_int),
            "mel_get_contacts": ([vp, ctypes.c_int, P(ctypes.c_int32), P(ctypes.c_double), ctypes.c_size_t,
                                  P(ctypes.c_size_t)], ctypes.c_int),
            "mel_get_neighbours": ([vp, ctypes.c_int, P(ctypes.c_int32), ctypes.c_size_t, P(ctypes.c_size_t)],
                                   ctypes.c_int),
            "mel_get_body_state": ([vp, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double)], ctypes.c_int),
            "mel_get_body_status": ([vp, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_double)], ctypes.c_int),
            "mel_set_monitoring": ([vp, ctypes.c_int], ctypes.c_int),
            "mel_get_monitoring": ([vp, P(ctypes.c_double), ctypes.c_size_t, P(ctypes.c_size_t)], ctypes.c_int),
            "mel_write_monitoring": ([vp, ctypes.c_char_p], ctypes.c_int),
            "mel_get_body_sizes": ([vp, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                    P(ctypes.c_int)], ctypes.c_int),
            "mel_compute_nodal_stresses": ([vp], ctypes.c_int),
            "mel_get_contact_internals": ([vp, ctypes.c_int, P(ctypes.c_double), ctypes.c_size_t, P(ctypes.c_size_t)],
                                          ctypes.c_int),
            "mel_write_graphic": ([vp, ctypes.c_char_p, P(ctypes.c_int), ctypes.c_int], ctypes.c_int),
            "mel_get_nodal_stress": ([vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_double), ctypes.c_size_t],
                                     ctypes.c_int),
            "mel_get_border_field": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_double),
                                      ctypes.c_size_t], ctypes.c_int),
            "mel_get_body_energy": ([vp, ctypes.c_int, P(ctypes.c_double)], ctypes.c_int),
            "mel_get_scalars": ([vp, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double)], ctypes.c_int),
            "mel_eval_forces": ([vp, ctypes.c_uint], ctypes.c_int),
            "mel_get_state_size": ([vp, P(ctypes.c_size_t)], ctypes.c_int),
            "mel_save_state": ([vp, vp, ctypes.c_size_t, P(ctypes.c_size_t)], ctypes.c_int),
            "mel_load_state": ([vp, vp, ctypes.c_size_t], ctypes.c_int),
            "mel_eval_forces_noalloc": ([vp, ctypes.c_uint], ctypes.c_int),
            "mel_get_info": ([vp, P(Info)], ctypes.c_int),
            "mel_write_dynamic": ([vp, ctypes.c_char_p], ctypes.c_int),
            "mel_get_stream": ([vp, P(vp)], ctypes.c_int),
            "mel_set_kernel_timing": ([vp, ctypes.c_int], ctypes.c_int),
            "mel_get_kernel_timing": ([vp, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_longlong)], ctypes.c_int),
            "mel_kernel_count": ([], ctypes.c_int),
            "mel_kernel_name": ([ctypes.c_int], ctypes.c_char_p),
            "mel_destroy": ([vp], None),
            "mel_last_error": ([vp], ctypes.c_char_p),
            "mel_version": ([], ctypes.c_char_p),
            "mel_get_device_bytes": ([vp, P(ctypes.c_size_t)], ctypes.c_int),
            "mel_set_graph_replay": ([vp, ctypes.c_int], ctypes.c_int),
            "mel_server_set_graph_replay": ([vp, ctypes.c_int], ctypes.c_int),
            "mel_server_create": ([ctypes.c_int, ctypes.c_size_t, ctypes.c_longlong, P(vp)], ctypes.c_int),
            "mel_server_submit": ([vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double,
                                   ctypes.c_int, P(ctypes.c_int)], ctypes.c_int),
            "mel_server_run": ([vp, ctypes.c_double, P(ctypes.c_int)], ctypes.c_int),
            "mel_server_job": ([vp, ctypes.c_int, P(JobInfo)], ctypes.c_int),
            "mel_server_job_context": ([vp, ctypes.c_int, P(vp)], ctypes.c_int),
            "mel_server_release": ([vp, ctypes.c_int], ctypes.c_int),
            "mel_server_usage": ([vp, P(ctypes.c_size_t), P(ctypes.c_size_t)], ctypes.c_int),
            "mel_server_job_error": ([vp, ctypes.c_int], ctypes.c_char_p),
            "mel_server_destroy": ([vp], None),
        }
        for name, (args, res) in sigs.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


def exported_symbols():
    return ["mel_load", "mel_load_binary", "mel_load_ex", "mel_load_binary_ex", "mel_default_options",
            "mel_get_options", "mel_autotune", "mel_update_proximity", "mel_euler_steps", "mel_synchronize",
            "mel_solver_steps", "mel_get_solver_info", "mel_get_body_error", "mel_get_mass_scaling",
            "mel_get_node_field", "mel_set_node_field", "mel_get_body_sizes", "mel_compute_nodal_stresses", "mel_write_graphic",
            "mel_get_contact_internals",
            "mel_get_nodal_stress", "mel_get_contacts", "mel_get_neighbours",
            "mel_get_body_state", "mel_get_body_status", "mel_set_monitoring", "mel_get_monitoring",
            "mel_write_monitoring", "mel_get_body_energy", "mel_get_scalars", "mel_eval_forces",
            "mel_eval_forces_noalloc", "mel_get_state_size", "mel_save_state", "mel_load_state",
            "mel_get_border_field", "mel_get_info", "mel_write_dynamic", "mel_get_stream",
            "mel_set_kernel_timing", "mel_get_kernel_timing", "mel_kernel_count", "mel_kernel_name",
            "mel_destroy", "mel_last_error", "mel_version", "mel_get_device_bytes", "mel_set_graph_replay",
            "mel_server_set_graph_replay",
            "mel_server_create", "mel_server_submit", "mel_server_run", "mel_server_job",
            "mel_server_job_context", "mel_server_release", "mel_server_usage", "mel_server_job_error",
            "mel_server_destroy"]


class Context:
    """One loaded model on one HIP device (owns the mel_ctx)."""

    def __init__(self, path, save_index=0, precision=FP64, device=0, binary=False, options=None):
        lib = library()
        h = ctypes.c_void_p()
        if isinstance(options, dict):
            options = make_options(**options)
        opt = ctypes.byref(options) if options is not None else None
        if binary:
            st = lib.mel_load_binary_ex(os.fsencode(path), precision, device, opt, ctypes.byref(h))
        else:
            st = lib.mel_load_ex(os.fsencode(path), save_index, precision, device, opt, ctypes.byref(h))
        if st != MEL_OK:
            raise MelError(st, lib.mel_last_error(None).decode())
        self._h = h
        self.lib = lib
        self._owner = None
        self.info = self.get_info()

    def options(self):
        """The launch configuration this context runs with (mel_get_options)."""
        o = Options()
        self._check(self.lib.mel_get_options(self._h, ctypes.byref(o)))
        return o.as_dict()

    @classmethod
    def _borrowed(cls, handle, owner):
        """A view of a context that `owner` (a Server) holds and destroys."""
        self = cls.__new__(cls)
        self._h = handle
        self.lib = library()
        self._owner = owner
        self.info = self.get_info()
        return self

    def _check(self, st):
        if st != MEL_OK:
            raise MelError(st, self.lib.mel_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owner", None) is None:
                self.lib.mel_destroy(self._h)
            self._h = None

    def set_graph_replay(self, on):
        self._check(self.lib.mel_set_graph_replay(self._h, int(bool(on))))

    def device_bytes(self):
        n = ctypes.c_size_t()
        self._check(self.lib.mel_get_device_bytes(self._h, ctypes.byref(n)))
        return n.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_info(self):
        inf = Info()
        self._check(self.lib.mel_get_info(self._h, ctypes.byref(inf)))
        return inf

    def euler_steps(self, n):
        self._check(self.lib.mel_euler_steps(self._h, n))

    def solver_steps(self, n):
        """n iterations with the model's scheme (mel_solver_steps)."""
        self._check(self.lib.mel_solver_steps(self._h, n))

    def solver_info(self):
        inf = SolverInfo()
        self._check(self.lib.mel_get_solver_info(self._h, ctypes.byref(inf)))
        return inf

    def body_error(self, body):
        """(total_error, max_error, node_for_max_error, mass_mass_scaling,
        factor_mass_scaling) of the last attempt (mel_get_body_error)."""
        out = np.zeros(5)
        self._check(self.lib.mel_get_body_error(self._h, body, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def mass_scaling(self, body, n):
        """Node mass-scaling factors (x, y) of a body (mel_get_mass_scaling)."""
        fx, fy = np.zeros(n), np.zeros(n)
        P = ctypes.POINTER(ctypes.c_double)
        self._check(self.lib.mel_get_mass_scaling(self._h, body, fx.ctypes.data_as(P), fy.ctypes.data_as(P), n))
        return fx, fy

    def update_proximity(self):
        self._check(self.lib.mel_update_proximity(self._h))

    def synchronize(self):
        self._check(self.lib.mel_synchronize(self._h))

    def stream(self):
        s = ctypes.c_void_p()
        self._check(self.lib.mel_get_stream(self._h, ctypes.byref(s)))
        return s.value

    def n_nodes(self, body):
        out = np.zeros(1)
        # size query: the field length equals the body's node count
        return self._body_nodes[body]

    def node_field(self, body, field, n):
        if isinstance(field, str):
            field = NODE_FIELDS.index(field)
        out = np.zeros(n, dtype=np.float64)
        self._check(self.lib.mel_get_node_field(self._h, body, field,
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        return out

    def set_node_field(self, body, field, values):
        if isinstance(field, str):
            field = NODE_FIELDS.index(field)
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self.lib.mel_set_node_field(self._h, body, field,
                                                v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(v)))

    def contacts(self, body):
        cnt = ctypes.c_size_t()
        self._check(self.lib.mel_get_contacts(self._h, body, None, None, 0, ctypes.byref(cnt)))
        n = cnt.value
        ids = np.zeros(16 * max(n, 1), dtype=np.int32)
        st = np.zeros(18 * max(n, 1), dtype=np.float64)
        self._check(self.lib.mel_get_contacts(self._h, body, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                              st.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                              ctypes.byref(cnt)))
        return ids[:16 * n], st[:18 * n]

    def neighbours(self, body):
        cnt = ctypes.c_size_t()
        self._check(self.lib.mel_get_neighbours(self._h, body, None, 0, ctypes.byref(cnt)))
        out = np.zeros(3 * max(cnt.value, 1), dtype=np.int32)
        self._check(self.lib.mel_get_neighbours(self._h, body, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                cnt.value, ctypes.byref(cnt)))
        return out[:3 * cnt.value]

    def nodal_stresses(self):
        """Body::Compute_nodal_stresses at the current state: one array of
        shape (len(STRESS_FIELDS), nodes) per body (mel_compute_nodal_stresses,
        mel_get_nodal_stress)."""
        self._check(self.lib.mel_compute_nodal_stresses(self._h))
        out = []
        for b in range(self.get_info().n_bodies):
            n = self.body_sizes(b)[0]
            a = np.zeros((len(STRESS_FIELDS), n))
            for f in range(len(STRESS_FIELDS)):
                self._check(self.lib.mel_get_nodal_stress(self._h, b, f, a[f].ctypes.data_as(
                    ctypes.POINTER(ctypes.c_double)), n))
            out.append(a)
        return out

    def contact_internals(self, body):
        """(elements, 3) internal variables of body's contact elements (mel_get_contact_internals)."""
        n = ctypes.c_size_t()
        self._check(self.lib.mel_get_contact_internals(self._h, body, None, 0, ctypes.byref(n)))
        out = np.zeros(3 * n.value)
        self._check(self.lib.mel_get_contact_internals(self._h, body, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                        n.value, ctypes.byref(n)))
        return out

    def write_graphic(self, path, to_plot=None, number_print=0):
        """Write_graphic's VTK file of the current state (mel_write_graphic);
        to_plot: 41 flags of the reference's To_Plot slots, or None."""
        arr = None
        if to_plot is not None:
            arr = (ctypes.c_int * 41)(*[int(bool(x)) for x in to_plot])
        self._check(self.lib.mel_write_graphic(self._h, os.fsencode(path), arr, int(number_print)))

    def body_sizes(self, body):
        """(nodes, Gauss points, borders, rigid) of a body (mel_get_body_sizes)."""
        v = [ctypes.c_int() for _ in range(4)]
        self._check(self.lib.mel_get_body_sizes(self._h, body, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def body_state(self, body):
        r = np.zeros(26)
        w = np.zeros(7)
        self._check(self.lib.mel_get_body_state(self._h, body, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return r, w

    def body_status(self, body):
        """(status, damage) of a body: 0 active, 1 inactive, 2 nogravity;
        Body::damage (mel_get_body_status)."""
        st = ctypes.c_int()
        dm = ctypes.c_double()
        self._check(self.lib.mel_get_body_status(self._h, body, ctypes.byref(st), ctypes.byref(dm)))
        return st.value, dm.value

    def set_monitoring(self, on=True):
        """Form a Monitoring row after every step (mel_set_monitoring)."""
        self._check(self.lib.mel_set_monitoring(self._h, 1 if on else 0))

    def monitoring(self):
        """The Monitoring rows formed since the last call, as a list of
        arrays (mel_get_monitoring)."""
        n = ctypes.c_size_t()
        self._check(self.lib.mel_get_monitoring(self._h, None, 0, ctypes.byref(n)))
        buf = np.zeros(max(1, n.value))
        self._check(self.lib.mel_get_monitoring(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                buf.size, ctypes.byref(n)))
        rows, i = [], 0
        while i < n.value:
            k = int(buf[i])
            rows.append(buf[i + 1:i + 1 + k].copy())
            i += 1 + k
        return rows

    def write_monitoring(self, path):
        """Append the rows formed since the last read to path (mel_write_monitoring)."""
        self._check(self.lib.mel_write_monitoring(self._h, os.fsencode(path)))

    def body_energy(self, body):
        """{kinetic, kinetic from velocity parameters, boundary work of the
        last step} of one body (mel_get_body_energy)."""
        e = np.zeros(3)
        self._check(self.lib.mel_get_body_energy(self._h, body, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return e

    def scalars(self):
        """(time, works summed over bodies [7], energy scalars summed over
        bodies [3]) (mel_get_scalars)."""
        t = ctypes.c_double()
        w = np.zeros(7)
        e = np.zeros(3)
        self._check(self.lib.mel_get_scalars(self._h, ctypes.byref(t), w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             e.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return t.value, w, e

    def eval_forces(self, what=EVAL_ALL):
        """One force evaluation at the current state (mel_eval_forces)."""
        self._check(self.lib.mel_eval_forces(self._h, int(what)))

    def save_state(self):
        """The dynamic state as bytes (mel_save_state)."""
        n = ctypes.c_size_t()
        self._check(self.lib.mel_get_state_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        w = ctypes.c_size_t()
        self._check(self.lib.mel_save_state(self._h, buf, n.value, ctypes.byref(w)))
        return buf.raw[:w.value]

    def load_state(self, blob):
        """Restore a state saved from a context of the same model (mel_load_state)."""
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        self._check(self.lib.mel_load_state(self._h, buf, len(blob)))

    def eval_forces_noalloc(self, what=EVAL_ALL):
        """mel_eval_forces without device allocation (mel_eval_forces_noalloc)."""
        self._check(self.lib.mel_eval_forces_noalloc(self._h, int(what)))

    def border_field(self, body, border, field, n):
        out = np.zeros(n)
        self._check(self.lib.mel_get_border_field(self._h, body, border, field,
                                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        return out

    def kernel_names(self):
        return [self.lib.mel_kernel_name(k).decode() for k in range(self.lib.mel_kernel_count())]

    def set_kernel_timing(self, kernels):
        """kernels: False/None (off), True (all) or an iterable of kernel names."""
        if kernels is True:
            mask = -1
        elif not kernels:
            mask = 0
        else:
            names = self.kernel_names()
            mask = 0
            for k in kernels:
                mask |= 1 << names.index(k)
        self._check(self.lib.mel_set_kernel_timing(self._h, mask))

    def kernel_timing(self):
        """{kernel name: (total ms, launches)} of the instrumented launches."""
        out = {}
        for k in range(self.lib.mel_kernel_count()):
            ms = ctypes.c_double()
            n = ctypes.c_longlong()
            self._check(self.lib.mel_get_kernel_timing(self._h, k, ctypes.byref(ms), ctypes.byref(n)))
            out[self.lib.mel_kernel_name(k).decode()] = (ms.value, n.value)
        return out

    def write_dynamic(self, path):
        self._check(self.lib.mel_write_dynamic(self._h, os.fsencode(path)))


class Server:
    """Job server (mel_server_*): independent jobs, each a binary model
    container run for a number of steps, admitted by device footprint under
    an HBM budget and run concurrently on one device."""

    def __init__(self, hbm_budget, steps_per_slice=50, device=0, graph_replay=None):
        lib = library()
        h = ctypes.c_void_p()
        st = lib.mel_server_create(device, int(hbm_budget), int(steps_per_slice), ctypes.byref(h))
        if st != MEL_OK:
            raise MelError(st, "mel_server_create")
        self._h = h
        self.lib = lib
        self._views = {}
        # None: each job keeps its context's own choice (MELODY2D_GRAPHS included)
        if graph_replay is not None:
            self._check(lib.mel_server_set_graph_replay(h, int(bool(graph_replay))))

    def _check(self, st, job=-1):
        if st != MEL_OK:
            raise MelError(st, self.lib.mel_server_job_error(self._h, job).decode())

    def submit(self, path, steps, precision=FP64, arrival=0.0, keep=False):
        j = ctypes.c_int()
        self._check(self.lib.mel_server_submit(self._h, os.fsencode(path), precision, int(steps), float(arrival),
                                               int(bool(keep)), ctypes.byref(j)))
        return j.value

    def run(self, timeout=3600.0):
        """Drive the loop until every job is done or failed; returns the
        number still unfinished when `timeout` seconds have passed."""
        left = ctypes.c_int()
        self._check(self.lib.mel_server_run(self._h, float(timeout), ctypes.byref(left)))
        return left.value

    def job(self, j):
        inf = JobInfo()
        self._check(self.lib.mel_server_job(self._h, j, ctypes.byref(inf)))
        return inf

    def job_error(self, j):
        return self.lib.mel_server_job_error(self._h, j).decode()

    def context(self, j):
        """The held context of a kept, done job (valid until release)."""
        if j not in self._views:
            h = ctypes.c_void_p()
            self._check(self.lib.mel_server_job_context(self._h, j, ctypes.byref(h)), j)
            self._views[j] = Context._borrowed(h, self)
        return self._views[j]

    def release(self, j):
        v = self._views.pop(j, None)
        if v is not None:
            v._h = None
        self._check(self.lib.mel_server_release(self._h, j), j)

    def usage(self):
        u, b = ctypes.c_size_t(), ctypes.c_size_t()
        self._check(self.lib.mel_server_usage(self._h, ctypes.byref(u), ctypes.byref(b)))
        return u.value, b.value

    def close(self):
        if getattr(self, "_h", None):
            for v in self._views.values():
                v._h = None
            self._views = {}
            self.lib.mel_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def autotune(binary_path, precision=FP64, device=0, steps=100):
    """Time every candidate launch configuration of a binary model (mel_autotune).

    Returns (options dict, ms per step, configurations timed)."""
    lib = library()
    best = Options()
    ms = ctypes.c_double()
    n = ctypes.c_int()
    st = lib.mel_autotune(os.fsencode(binary_path), precision, device, steps, ctypes.byref(best),
                          ctypes.byref(ms), ctypes.byref(n))
    if st != MEL_OK:
        raise MelError(st, lib.mel_last_error(None).decode())
    return best.as_dict(), ms.value, n.value
