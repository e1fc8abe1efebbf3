"""The allocation-free face-flux pass, sfx_face_flux_pass_static: the face
fluxes live in the library's static device scratch instead of a caller
workspace.  Two calls in a row on one device, the second on a smaller mesh,
each equal bit for bit to sfx_face_flux_pass with a workspace (the same
kernels run); a mesh whose fluxes exceed the scratch is refused before
anything is launched."""
import ctypes as ct

import numpy as np
import pytest

from shadowfax_amd import _capi as C
from shadowfax_amd import mesh as M
from shadowfax_amd.fluxpass import FluxPass
from shadowfax_amd.riemann import RiemannSolver

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def bundle(nd, n, kind):
    if nd == 2:
        m = M.voronoi_periodic(M.hilbert_sorted(M.jittered_grid(n, 0.25, seed=n)))
        return M.prepare(m, M.kh_ics(m), G, kind=kind)
    m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(n, seed=n)))
    return M.prepare(m, M.sedov_ics(m, ncells=8), G, kind=kind)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("nd,kind,sizes", [(2, 1, (96, 40)), (3, 0, (6000, 2500))])
def test_static_pass_equals_workspace_pass(device, prec, nd, kind, sizes):
    import torch
    s = RiemannSolver(kind, nd, G, 1e-8, -5.0, prec)
    for n in sizes:  # the second call: a smaller mesh on the same scratch
        print("case %s %dD kind %d n %d" % (prec, nd, kind, n), flush=True)
        b = bundle(nd, n, kind)
        ref = FluxPass(s, M.DeviceMesh(b, prec, device))
        ref.run()
        st = FluxPass(s, M.DeviceMesh(b, prec, device))
        st.m.dQ.fill_(np.nan)
        st.run_static()
        torch.cuda.synchronize()
        assert torch.equal(st.m.dQ, ref.m.dQ)
        assert torch.equal(st.m.grad, ref.m.grad)
        assert torch.equal(st.m.fstatus, ref.m.fstatus)
        assert int((st.m.dQL != 0).sum()) == 0  # the mesh's own buffer is not used
        assert float(ref.m.dQ.abs().max()) > 0


def test_static_pass_refuses_meshes_beyond_the_scratch(device):
    s = RiemannSolver(1, 2, G, 1e-8, -5.0, "mixed")
    b = bundle(2, 16, 1)
    fp = FluxPass(s, M.DeviceMesh(b, "mixed", device))
    d = C.MeshDesc.from_buffer_copy(fp.m.desc)
    d.F = C.SFX_STATIC_SCRATCH_BYTES // (4 * 4) + 1  # one face more than fits (4 FP32 values)
    rc = C.lib().sfx_face_flux_pass_static(s.handle, ct.byref(d), ct.c_void_p(fp.m.grad.data_ptr()),
                                           ct.c_void_p(fp.m.dQ.data_ptr()), None, None)
    assert rc == C.SFX_ERR_ARG
    assert b"static scratch" in C.lib().sfx_last_error()
