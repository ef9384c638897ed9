"""Model / Body / Solver façade with the reference's call structure.

Code/main.cpp loads a model (Load_static, Load_dynamic and the
initialisation block of Code/main.cpp:83-126), then loops Euler_step
(Code/Solver.h:10-88) with Update_proximity every Contact_update_period
(Code/main.cpp:191-196) and Write_dynamic at every save (Code/IO.h:2351).
The same structure over the C-ABI (melody2d_amd.native), as
include/melody2d.hpp gives it to C++ callers:

    model = Model.load("run_dir", number_save=0)
    solver = Solver()
    while model.time < tend:
        solver.euler_step(model)
    model.write_dynamic("DYNAMIC_00001.asc")

Every call goes to the HIP library; there is no CPU path.
"""
from . import native


class Body:
    """One body of a loaded model (Code/Body.h): its node fields, contact
    elements and body-level state, read from the device on each call."""

    def __init__(self, ctx, index):
        self._ctx = ctx
        self.index = index
        self.nb_nodes, self.nb_gauss, self.nb_borders, rigid = ctx.body_sizes(index)
        self.rigid = bool(rigid)

    def field(self, name):
        """A node field (native.NODE_FIELDS name) of every node of the body."""
        return self._ctx.node_field(self.index, name, self.nb_nodes)

    def set_field(self, name, values):
        self._ctx.set_node_field(self.index, name, values)

    def contacts(self):
        """(ids, state) of the contact elements with this body as slave."""
        return self._ctx.contacts(self.index)

    def state(self):
        """(rigid record of 26 values, the 7 works)."""
        return self._ctx.body_state(self.index)

    def energy(self):
        return self._ctx.body_energy(self.index)

    def status(self):
        """(Body::status: 0 active, 1 inactive, 2 nogravity; Body::damage)."""
        return self._ctx.body_status(self.index)


class Model:
    """A model on one HIP device (owns its context)."""

    def __init__(self, ctx):
        self._ctx = ctx
        self.bodies = [Body(ctx, b) for b in range(ctx.get_info().n_bodies)]

    @classmethod
    def load(cls, directory, number_save=0, precision=native.FP64, device=0, options=None):
        """Load_static + Load_dynamic (Code/IO.h:598, :1871) of the model
        files in `directory` and the initialisation block of Code/main.cpp."""
        return cls(native.Context(directory, save_index=number_save, precision=precision, device=device,
                                  options=options))

    @classmethod
    def load_binary(cls, path, precision=native.FP64, device=0, options=None):
        return cls(native.Context(path, binary=True, precision=precision, device=device, options=options))

    @property
    def context(self):
        return self._ctx

    @property
    def time(self):
        return self._ctx.get_info().time

    @property
    def deltat(self):
        return self._ctx.get_info().dt

    @property
    def number_iteration(self):
        return self._ctx.get_info().iteration

    def scalars(self):
        """(time, works summed over bodies, kinetic energies) of the
        Monitoring energy block (Code/Monitoring.h:508-537)."""
        return self._ctx.scalars()

    def write_dynamic(self, path):
        """Write_dynamic (Code/IO.h:2351-2636)."""
        self._ctx.write_dynamic(path)

    def start_monitoring(self):
        """Monitoring rows after every step from now on (Code/Monitoring.h:8-559)."""
        self._ctx.set_monitoring(True)

    def write_monitoring(self, path):
        """Write_monitoring (Code/IO.h:2644-2666): the rows since the last write, appended."""
        self._ctx.write_monitoring(path)

    def close(self):
        self._ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Solver:
    """The explicit solver: Euler_step with the contact-update cadence."""

    def euler_step(self, model, n=1):
        """n Euler_steps (Code/Solver.h:10-88), each followed by
        Update_proximity when the contact-update period has elapsed
        (Code/main.cpp:191-196)."""
        model.context.euler_steps(n)

    def solver_step(self, model, n=1):
        """n iterations with the model's scheme: Euler_step, or Solver_step
        (Code/Solver.h:98-486) for the adaptive schemes, each followed by the
        contact-update cadence (Code/main.cpp:148-196)."""
        model.context.solver_steps(n)

    def update_proximity(self, model):
        """Update_proximity (Code/Proximity.h:48-473) now."""
        model.context.update_proximity()

    def eval_forces(self, model, what=native.EVAL_ALL):
        """One force evaluation without integration."""
        model.context.eval_forces(what)
