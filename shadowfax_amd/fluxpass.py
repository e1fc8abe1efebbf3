"""The face-flux pass on the GPU, mirroring the reference's hydro step:

  VorTessManager::estimate_gradients -> VorCell::estimate_gradient (per cell)
  VorTessManager::hydro -> VorFace::set_v + VorFace::calculate_flux (per face)
  GasParticle::increase_dQ in face order, then update_Q + get_W (per cell)

(Shadowfax src/Simulation.cpp:414-483, src/VorTessManager.cpp:198-296).
Everything runs through the C ABI (sfx_gradients, sfx_face_fluxes,
sfx_gather_dQ, sfx_update_step) on a DeviceMesh.
"""
import ctypes as ct

from . import _capi as C
from .mesh import DeviceMesh
from .riemann import RiemannSolver, _stream


class FluxPass:
    def __init__(self, solver: RiemannSolver, dmesh: DeviceMesh):
        if (solver.precision == "mixed") != (dmesh.storage == "mixed"):
            raise ValueError("solver precision and mesh storage must match")
        if solver.ndim != dmesh.ndim:
            raise ValueError("solver ndim %d != mesh ndim %d" % (solver.ndim, dmesh.ndim))
        self.s = solver
        self.m = dmesh
        self._d = ct.byref(dmesh.desc)
        if not dmesh.egeo_built:
            dmesh.build_entries(solver.handle)

    def _v(self, t):
        return ct.c_void_p(t.data_ptr())

    def gradients(self, stream=None):
        C.check(C.lib().sfx_gradients(self.s.handle, self._d, self._v(self.m.grad),
                                      _stream(stream)), "sfx_gradients")

    def face_fluxes(self, stream=None, status=True):
        st = self._v(self.m.fstatus) if status else None
        C.check(C.lib().sfx_face_fluxes(self.s.handle, self._d, self._v(self.m.grad),
                                        self._v(self.m.dQL), st, _stream(stream)),
                "sfx_face_fluxes")

    def gather(self, stream=None):
        C.check(C.lib().sfx_gather_dQ(self.s.handle, self._d, self._v(self.m.dQL),
                                      self._v(self.m.dQ), _stream(stream)), "sfx_gather_dQ")

    def run(self, stream=None, status=True):
        """Gradients + face fluxes + deterministic gather into dmesh.dQ."""
        st = self._v(self.m.fstatus) if status else None
        C.check(C.lib().sfx_face_flux_pass(self.s.handle, self._d, self._v(self.m.grad),
                                           self._v(self.m.dQL), self._v(self.m.dQ), st,
                                           _stream(stream)), "sfx_face_flux_pass")

    def run_static(self, stream=None, status=True):
        """run() without a face-flux workspace: sfx_face_flux_pass_static
        keeps the face fluxes in the library's static device scratch (for
        callers that cannot allocate; calls that overlap on one device race
        on it).  dmesh.dQL is not written."""
        st = self._v(self.m.fstatus) if status else None
        C.check(C.lib().sfx_face_flux_pass_static(self.s.handle, self._d, self._v(self.m.grad),
                                                  self._v(self.m.dQ), st, _stream(stream)),
                "sfx_face_flux_pass_static")

    def update(self, stream=None, endtime=None, current_time=0):
        """Q -= dQ, dQ = 0, W = get_W(V, Q) (writes the mesh's W in place),
        and the mesh's masses take the new Q.m (read by the next pass's
        mass-flux limiter).  With endtime (a device uint64 tensor [C]) W is
        rebuilt only where endtime == current_time, the time after the step
        (individual time steps, src/Simulation.cpp:792-796)."""
        a = self.m.arr
        et = None if endtime is None else self._v(endtime)
        C.check(C.lib().sfx_update_step_timed(self.s.handle, self.m.C, self._v(self.m.Q),
                                              self._v(self.m.dQ), self._v(a["volume"]),
                                              self._v(a["W"]), self._v(a["mass"]), et,
                                              int(current_time), self._v(self.m.cstatus),
                                              _stream(stream)),
                "sfx_update_step_timed")

    def step(self, stream=None):
        self.run(stream)
        self.update(stream)

