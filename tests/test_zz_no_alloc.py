"""The allocation-free reaction potential (ibim_polarize_device_noalloc)
against ibim_polarize_device, through the C ABI: two calls in a row on
different plans, the second with fewer targets on a smaller surface; FP64
bar 1e-10 relative (the two differ in the far field's summation order: the
no-alloc form evaluates the V list directly instead of by GEMMs).  Also: a
request too large for the scratch is rejected with an error, and device
memory in use does not grow across calls."""
import ctypes as C

import numpy as np
import pytest

import geometry as G

pytestmark = pytest.mark.gpu

PB = dict(kappa=0.1257, eps_in=2.0, eps_out=80.0, vacant_radius=0.05)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def run_pair(ibim, torch, dev, pts, nrm, w, atoms, seed, p=4):
    x = np.random.default_rng(seed).uniform(-1, 1, 2 * len(pts))
    d_at = torch.from_numpy(np.ascontiguousarray(atoms)).to(dev)
    d_x = torch.from_numpy(x).to(dev)
    a = torch.zeros(len(atoms), dtype=torch.float64, device=dev)
    b = torch.full((len(atoms),), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    P = lambda t: C.c_void_p(t.data_ptr())
    with ibim.Plan(pts, nrm, w, p=p, max_points=32, max_level=10, **PB) as plan:
        plan.polarize_device(len(atoms), P(d_at), P(d_x), P(a))
        plan.sync()
        free0 = torch.cuda.mem_get_info()[0]
        plan.polarize_device_noalloc(len(atoms), P(d_at), P(d_x), P(b))
        free1 = torch.cuda.mem_get_info()[0]
    return a.cpu().numpy(), b.cpu().numpy(), free0 - free1


def test_noalloc_matches_polarize_two_calls(ibim, gpu):
    import torch
    rng = np.random.default_rng(11)
    pts, nrm, w = G.shell(4000, seed=21)
    atoms = np.concatenate([rng.uniform(-1.3, 1.3, size=(300, 3)), [[3.0, 0.1, 0.2]]])
    a, b, grew = run_pair(ibim, torch, gpu, pts, nrm, w, atoms, 1)
    print(f"call 1: {len(atoms)} targets, rel diff {rel(b, a):.2e}")
    assert np.all(np.isfinite(b)) and rel(b, a) < 1e-10
    assert grew <= 0, f"device memory in use grew by {grew} bytes"
    # second call: smaller surface, fewer targets (the array keeps the first
    # call's contents; the result must not depend on them)
    pts2, nrm2, w2 = G.clusters(1500, seed=22)
    atoms2 = rng.uniform(-1.5, 1.5, size=(17, 3))
    a2, b2, grew2 = run_pair(ibim, torch, gpu, pts2, nrm2, w2, atoms2, 2)
    print(f"call 2: {len(atoms2)} targets, rel diff {rel(b2, a2):.2e}")
    assert np.all(np.isfinite(b2)) and rel(b2, a2) < 1e-10
    assert grew2 <= 0


def test_noalloc_rejects_beyond_its_scratch(ibim, gpu):
    import torch
    pts, nrm, w = G.shell(500, seed=3)
    m = 1_000_000  # about 104 MB of target scratch > 64 MiB
    d_at = torch.zeros((m, 3), dtype=torch.float64, device=gpu)
    d_x = torch.zeros(2 * len(pts), dtype=torch.float64, device=gpu)
    d_out = torch.zeros(m, dtype=torch.float64, device=gpu)
    P = lambda t: C.c_void_p(t.data_ptr())
    with ibim.Plan(pts, nrm, w, p=4, max_points=32, max_level=10, **PB) as plan:
        with pytest.raises(ibim.IbimError, match="MiB"):
            plan.polarize_device_noalloc(m, P(d_at), P(d_x), P(d_out))
