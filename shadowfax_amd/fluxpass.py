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

    def update(self, stream=None):
        """Q -= dQ, dQ = 0, W = get_W(V, Q) (writes the mesh's W in place)."""
        a = self.m.arr
        C.check(C.lib().sfx_update_step(self.s.handle, self.m.C, self._v(self.m.Q),
                                        self._v(self.m.dQ), self._v(a["volume"]),
                                        self._v(a["W"]), self._v(self.m.cstatus),
                                        _stream(stream)), "sfx_update_step")

    def step(self, stream=None):
        self.run(stream)
        self.update(stream)


def smoke(dev):
    """A 16x16 periodic Kelvin-Helmholtz mesh: FP64 flux pass vs the oracle,
    mixed pass vs FP64."""
    import os
    import sys

    import numpy as np
    import torch

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, os.path.join(root, "oracle"))
    import oracle as O

    from . import mesh as M

    g = 5.0 / 3.0
    m = M.voronoi_periodic(M.jittered_grid(16, 0.25, seed=7))
    W = M.kh_ics(m)
    b = M.prepare(m, W, g, kind=1)
    s64 = RiemannSolver(1, 2, g, 1e-8, -5.0, "fp64")
    fp = FluxPass(s64, DeviceMesh(b, "fp64", dev))
    fp.run()
    torch.cuda.synchronize()
    dQ = fp.m.dQ_soa()
    ob = dict(b)
    ob["grad"] = O.gradients(1, b)
    dQo, _, _, _ = O.face_flux_pass(1, ob)
    scale = np.abs(dQo).max()
    err = np.abs(dQ - dQo).max() / scale
    print("smoke flux pass fp64 vs oracle: max err %.3e" % err, flush=True)
    assert err <= 1e-12
    smx = RiemannSolver(1, 2, g, 1e-8, -5.0, "mixed")
    fm = FluxPass(smx, DeviceMesh(b, "mixed", dev))
    fm.run()
    torch.cuda.synchronize()
    errm = np.abs(fm.m.dQ_soa() - dQ).max() / scale
    print("smoke flux pass mixed vs fp64: max err %.3e" % errm, flush=True)
    assert errm <= 1e-6
