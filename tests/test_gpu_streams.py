"""Stream-ordered calls (ibim_plan_set_stream): with the plan on torch's
stream, a matvec enqueued right after a long kernel that writes its input
sees that input, and later work on the stream sees the matvec's result, with
no host synchronisation in between.  A plan that ignored the stream would
read the input before the delayed write (the sleep kernel makes that race
certain) and fail the comparison."""
import ctypes as C

import numpy as np
import pytest

import geometry as G

pytestmark = pytest.mark.gpu

PB = dict(kappa=0.1257, eps_in=2.0, eps_out=80.0, vacant_radius=0.05)


@pytest.mark.parametrize("precision", ["fp64", "mixed16"])
def test_matvec_ordered_on_callers_stream(ibim, gpu, precision):
    import torch
    prec = {"fp64": ibim.PREC_FP64, "mixed16": ibim.PREC_MIXED_F16}[precision]
    pts, nrm, w = G.shell(3000, seed=8)
    xh = np.random.default_rng(9).uniform(-1, 1, 2 * len(pts))
    P = lambda t: C.c_void_p(t.data_ptr())
    with ibim.Plan(pts, nrm, w, p=4, max_points=32, max_level=10, precision=prec, **PB) as plan:
        ref = plan.matvec(xh)
        s = torch.cuda.Stream()
        x_src = torch.from_numpy(xh).to(gpu)
        x = torch.zeros_like(x_src)
        y = torch.zeros_like(x_src)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.set_stream(s.cuda_stream)
            torch.cuda._sleep(200_000_000)  # ~0.1 s of GPU time before x is written
            x.copy_(x_src)
            plan.matvec_device(P(x), P(y))
            z = y * 2.0
        plan.set_stream(None)
        s.synchronize()
        got = (z / 2.0).cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"{precision}: stream-ordered matvec vs host matvec rel {err:.2e}")
    assert err < 1e-12
