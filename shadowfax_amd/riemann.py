"""Python mirror of the reference's Riemann solver plugin interface
(Shadowfax src/riemann/RiemannSolver.hpp:37-155, RiemannSolverFactory.hpp),
backed by the HIP library through the C ABI.

Same names and argument meaning as the reference:
  ExactRiemannSolver(gamma=1.66667, tolerance=1e-8, cutoff=-5.)
  HLLCRiemannSolver(gamma)
  RiemannSolverFactory.generate(type, gamma, tolerance, cutoff)
  solver.solve(WL, WR, n, mach, reflective) -> (Whalf, mach)   (WR is
      overwritten by WL when reflective, as in the reference)
  solver.solve_for_flux(WL, WR, n, v, reflective) -> flux
  solver.tag(), solver.get_gamma(), solver.get_neval()
plus the batched entry point used by the flux pass and the benchmarks,
solve_for_flux_batch(), which takes device tensors in SoA layout.

Errors that the reference throws (RiemannSolverException) come back as
status bits; the scalar methods raise RiemannSolverException when the bit is
set, mirroring the reference's behaviour.
"""
import ctypes as ct

import numpy as np

from . import _capi as C

PRECISIONS = {"fp64": C.SFX_PREC_FP64, "mixed": C.SFX_PREC_MIXED}


class RiemannSolverException(RuntimeError):
    """src/riemann/RiemannSolver.hpp: thrown for a NaN/negative solution."""


def _ptr(t, dtype, shape, name):
    import torch
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor on the GPU" % name)
    if t.device.type != "cuda":
        raise ValueError("%s must be a device tensor" % name)
    if t.dtype != dtype:
        raise TypeError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if tuple(t.shape) != tuple(shape):
        raise ValueError("%s must have shape %s, got %s" % (name, shape, tuple(t.shape)))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    return ct.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    if stream is None:
        return ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ct.c_void_p(getattr(stream, "cuda_stream", stream))


class RiemannSolver:
    def __init__(self, kind, ndim=3, gamma=5.0 / 3.0, tolerance=1e-8, cutoff=-5.0,
                 precision="fp64"):
        self.kind = kind
        self.ndim = ndim
        self.precision = precision
        h = ct.c_void_p()
        C.check(C.lib().sfx_solver_create(kind, ndim, gamma, tolerance, cutoff,
                                          PRECISIONS[precision], ct.byref(h)),
                "sfx_solver_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            C.lib().sfx_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- reference accessors
    def tag(self):
        return C.lib().sfx_solver_tag(self._h).decode()

    def get_gamma(self):
        return C.lib().sfx_solver_gamma(self._h)

    def stats(self):
        s = C.Stats()
        C.check(C.lib().sfx_solver_get_stats(self._h, ct.byref(s)), "get_stats")
        return {k: getattr(s, k) for k, _ in C.Stats._fields_}

    def get_neval(self):
        return self.stats()["neval"]

    def reset_stats(self):
        C.check(C.lib().sfx_solver_reset_stats(self._h), "reset_stats")

    def enable_stats(self, on=True):
        C.check(C.lib().sfx_solver_enable_stats(self._h, 1 if on else 0), "enable_stats")

    # --- scalar drop-ins (host arrays)
    def _vec(self, x, n):
        a = np.zeros(n)
        if x is not None:
            a[:] = np.asarray(x, dtype=np.float64)[:n]
        return a

    def solve(self, WL, WR, n, mach=0.0, reflective=False):
        NC = self.ndim + 2
        wl = self._vec(WL, NC)
        wr = self._vec(WR, NC)
        nn = self._vec(n, self.ndim)
        out = np.zeros(NC)
        m = ct.c_double(mach)
        st = ct.c_uint8(0)
        dp = ct.POINTER(ct.c_double)
        C.check(C.lib().sfx_solve(self._h, wl.ctypes.data_as(dp), wr.ctypes.data_as(dp),
                                  nn.ctypes.data_as(dp), ct.byref(m), int(bool(reflective)),
                                  out.ctypes.data_as(dp), ct.byref(st)), "sfx_solve")
        if isinstance(WR, np.ndarray) and reflective:
            WR[:NC] = wr
        if st.value & C.ST_SOLVER:
            raise RiemannSolverException("Riemann solver returned an unphysical state")
        return out, m.value

    def solve_for_flux(self, WL, WR, n, v=None, reflective=False):
        NC = self.ndim + 2
        wl = self._vec(WL, NC)
        wr = self._vec(WR, NC)
        nn = self._vec(n, self.ndim)
        vv = self._vec(v, self.ndim)
        out = np.zeros(NC)
        st = ct.c_uint8(0)
        dp = ct.POINTER(ct.c_double)
        C.check(C.lib().sfx_solve_for_flux(self._h, wl.ctypes.data_as(dp),
                                           wr.ctypes.data_as(dp), nn.ctypes.data_as(dp),
                                           vv.ctypes.data_as(dp), int(bool(reflective)),
                                           out.ctypes.data_as(dp), ct.byref(st)),
                "sfx_solve_for_flux")
        if isinstance(WR, np.ndarray) and reflective:
            WR[:NC] = wl
        if st.value & C.ST_SOLVER:
            raise RiemannSolverException("Riemann solver returned an unphysical state")
        return out

    # --- batched (device tensors, SoA)
    def solve_for_flux_batch(self, WL, WR, n, v=None, reflective=None, F=None, status=None,
                             stream=None):
        """WL, WR: (ndim+2, N); n, v: (ndim, N); reflective, status: (N,) uint8.
        float64 tensors use the handle's precision; float32 tensors always use
        the mixed kernels.  Returns F (allocated if not given)."""
        import torch
        NC, ND = self.ndim + 2, self.ndim
        N = WL.shape[1]
        dt = WL.dtype
        if F is None:
            F = torch.empty((NC, N), dtype=dt, device=WL.device)
        args = [_ptr(WL, dt, (NC, N), "WL"), _ptr(WR, dt, (NC, N), "WR"),
                _ptr(n, dt, (ND, N), "n"), _ptr(v, dt, (ND, N), "v"),
                _ptr(reflective, torch.uint8, (N,), "reflective"),
                _ptr(F, dt, (NC, N), "F"), _ptr(status, torch.uint8, (N,), "status"),
                _stream(stream)]
        if dt == torch.float64:
            rc = C.lib().sfx_solve_for_flux_batch(self._h, N, *args)
        elif dt == torch.float32:
            rc = C.lib().sfx_solve_for_flux_batch_f32(self._h, N, *args)
        else:
            raise TypeError("states must be float32 or float64")
        C.check(rc, "solve_for_flux_batch")
        return F

    def solve_batch(self, WL, WR, n, reflective=None, Whalf=None, mach=None, status=None,
                    stream=None):
        import torch
        NC, ND = self.ndim + 2, self.ndim
        N = WL.shape[1]
        if Whalf is None:
            Whalf = torch.empty((NC, N), dtype=torch.float64, device=WL.device)
        f64 = torch.float64
        C.check(C.lib().sfx_exact_solve_batch(
            self._h, N, _ptr(WL, f64, (NC, N), "WL"), _ptr(WR, f64, (NC, N), "WR"),
            _ptr(n, f64, (ND, N), "n"), _ptr(reflective, torch.uint8, (N,), "reflective"),
            _ptr(Whalf, f64, (NC, N), "Whalf"), _ptr(mach, f64, (N,), "mach"),
            _ptr(status, torch.uint8, (N,), "status"), _stream(stream)), "solve_batch")
        return Whalf


    def set_split(self, iterations):
        """FP32 Newton iterations before a slow problem of a batched mixed
        exact solve is queued for the second, compacted kernel (0: one
        kernel).  Results do not depend on it."""
        C.check(C.lib().sfx_solver_set_split(self._h, int(iterations)), "set_split")

    def get_split(self):
        return C.lib().sfx_solver_get_split(self._h)

    # ---- per-cell operations over batches (FP64, device SoA) ----------

    def get_time_derivative_batch(self, W, grad, out=None, stream=None):
        """RiemannSolver::get_time_derivative for every column of W
        ([ndim+2][N]) with gradients grad ([ndim][ndim+2][N])."""
        import torch
        NC, ND = self.ndim + 2, self.ndim
        N = W.shape[1]
        f64 = torch.float64
        if out is None:
            out = torch.empty((NC, N), dtype=f64, device=W.device)
        C.check(C.lib().sfx_time_derivative_batch(
            self._h, N, _ptr(W, f64, (NC, N), "W"), _ptr(grad, f64, (ND, NC, N), "grad"),
            _ptr(out, f64, (NC, N), "dWdt"), _stream(stream)), "get_time_derivative_batch")
        return out

    def get_W_batch(self, volume, Q, W=None, stream=None):
        """RiemannSolver::get_W(volume, Q, true) per column; Q is updated in
        place where the pressure floor resets the energy."""
        import torch
        NC = self.ndim + 2
        N = Q.shape[1]
        f64 = torch.float64
        if W is None:
            W = torch.empty((NC, N), dtype=f64, device=Q.device)
        C.check(C.lib().sfx_get_W_batch(
            self._h, N, _ptr(volume, f64, (N,), "volume"), _ptr(Q, f64, (NC, N), "Q"),
            _ptr(W, f64, (NC, N), "W"), _stream(stream)), "get_W_batch")
        return W

    def hydro_timestep(self, W, vel, volume, timeline, starttime, endtime, dt=None,
                       active=None, status=None, result=None, stream=None):
        """TimeLine::calculate_timestep for gas cells (src/TimeLine.cpp:233-320):
        W [C][ndim+2], vel [C][ndim], volume [C] (FP64, cell-major);
        starttime/endtime [C] int64 tensors holding the unsigned times,
        advanced in place; timeline: dict (or _capi.Timeline) of integer_maxtime,
        current_time, timestep, max_timestep, min_timestep, tottime, mintime,
        cfl, global_timestep.  Returns (dt, active, status, result) with
        result = device int64[2]: {system step, number of active cells}."""
        import torch
        nd, NC = self.ndim, self.ndim + 2
        Cn = volume.shape[0]
        f64, dev = torch.float64, volume.device
        if dt is None:
            dt = torch.empty(Cn, dtype=f64, device=dev)
        if active is None:
            active = torch.empty(Cn, dtype=torch.uint8, device=dev)
        if status is None:
            status = torch.empty(Cn, dtype=torch.uint8, device=dev)
        if result is None:
            result = torch.empty(2, dtype=torch.int64, device=dev)
        tl = timeline
        if isinstance(timeline, dict):
            tl = C.Timeline()
            for k, v in timeline.items():
                setattr(tl, k, v)
        i64 = torch.int64
        C.check(C.lib().sfx_hydro_timestep(
            self._h, Cn, _ptr(W, f64, (Cn, NC), "W"), _ptr(vel, f64, (Cn, nd), "vel"),
            _ptr(volume, f64, (Cn,), "volume"), ct.byref(tl),
            _ptr(starttime, i64, (Cn,), "starttime"), _ptr(endtime, i64, (Cn,), "endtime"),
            _ptr(dt, f64, (Cn,), "dt"), _ptr(active, torch.uint8, (Cn,), "active"),
            _ptr(result, i64, (2,), "result"), _ptr(status, torch.uint8, (Cn,), "status"),
            _stream(stream)), "hydro_timestep")
        return dt, active, status, result

    def get_Q_batch(self, volume, W, Q=None, stream=None):
        """RiemannSolver::get_Q(volume, W) per column."""
        import torch
        NC = self.ndim + 2
        N = W.shape[1]
        f64 = torch.float64
        if Q is None:
            Q = torch.empty((NC, N), dtype=f64, device=W.device)
        C.check(C.lib().sfx_get_Q_batch(
            self._h, N, _ptr(volume, f64, (N,), "volume"), _ptr(W, f64, (NC, N), "W"),
            _ptr(Q, f64, (NC, N), "Q"), _stream(stream)), "get_Q_batch")
        return Q


class ExactRiemannSolver(RiemannSolver):
    """src/riemann/ExactRiemannSolver.hpp:107: defaults gamma=1.66667,
    tolerance=1e-8, cutoff=-5."""

    def __init__(self, gamma=1.66667, tolerance=1.e-8, cutoff=-5., ndim=3, precision="fp64"):
        super().__init__(C.SFX_EXACT, ndim, gamma, tolerance, cutoff, precision)


class HLLCRiemannSolver(RiemannSolver):
    """src/riemann/HLLCRiemannSolver.hpp:49."""

    def __init__(self, gamma, ndim=3, precision="fp64"):
        super().__init__(C.SFX_HLLC, ndim, gamma, 1e-8, -5.0, precision)

    def solve(self, *a, **k):
        raise NotImplementedError("HLLCRiemannSolver::solve aborts in the reference "
                                  "(src/riemann/HLLCRiemannSolver.cpp:240-245)")


class RiemannSolverFactory:
    """src/riemann/RiemannSolverFactory.hpp:115-155."""

    @staticmethod
    def generate(type="Exact", gamma=5.0 / 3.0, tolerance=1e-8, cutoff=-5.0, ndim=3,
                 precision="fp64"):
        if type == "Exact":
            return ExactRiemannSolver(gamma, tolerance, cutoff, ndim, precision)
        if type == "HLLC":
            return HLLCRiemannSolver(gamma, ndim, precision)
        raise C.SfxError("solver type %s has no GPU implementation" % type)

    @staticmethod
    def type(tag):
        return {"EXAC": "Exact", "HLLC": "HLLC"}[tag]
