"""The entry point that allocates nothing (abm_vf_state_variables_noalloc)
against abm_vf_state_variables, through the C ABI: two calls in a row, the
second smaller, on the same inputs to both entry points; dV/dphi identical,
the six float64 sums within 1e-12 relative (the tables' cos and sin come
from the device's libm in one and the host's in the other).  Sizes beyond
the static table space are rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, name, R, n, bits, vel, a0, b0, v0, prm, out, dphi):
    from abm_amd import _lib as L
    fn = getattr(lib, name)
    return fn(R, n, L.ptr(bits), L.ptr(vel), L.ptr(a0), L.ptr(b0), L.ptr(v0), prm[0], prm[1],
              prm[2], L.ptr(out), L.ptr(dphi), L.stream_ptr())


def _inputs(rng, R, n, dev):
    import torch
    from abm_amd.vf import pack_bits
    V = np.zeros((n, R), dtype=bool)
    for i in range(n):
        V[i] = rng.rand(R) < [0.0, 0.05, 0.5, 1.0][i % 4]
    V[4 % n, : R // 3] = True                      # one long run through ray 0
    bits = torch.as_tensor(pack_bits(V).view(np.int32).reshape(-1).copy(), device=dev)
    col = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
    return (bits, col(rng.uniform(-1, 2, n)), col(rng.uniform(0.1, 3, n)),
            col(rng.uniform(0.1, 3, n)), col(rng.uniform(0.5, 2, n)))


def test_no_alloc_state_variables_match_the_allocating_entry_point(gpu):
    import torch
    from abm_amd import _lib as L
    lib = L.load()
    rng = np.random.RandomState(12)
    prm = (0.1, 0.09, 0.09)
    for R, n in ((2048, 64), (600, 9)):           # the second call is smaller
        print(f"no-alloc state variables: R={R} n={n}", flush=True)
        ins = _inputs(rng, R, n, gpu)
        res = []
        for name in ("abm_vf_state_variables", "abm_vf_state_variables_noalloc"):
            out = torch.full((6 * n,), np.nan, dtype=torch.float64, device=gpu)
            dphi = torch.full((n * R,), 7, dtype=torch.int8, device=gpu)
            L.check(_call(lib, name, R, n, *ins, prm, out, dphi))
            torch.cuda.synchronize()
            res.append((out.cpu().numpy(), dphi.cpu().numpy()))
        (o1, d1), (o2, d2) = res
        assert np.isfinite(o1).all() and np.array_equal(d1, d2)
        assert np.allclose(o2, o1, rtol=1e-12, atol=1e-14), np.max(np.abs(o2 - o1))


def test_no_alloc_rejects_rays_beyond_its_table_space(gpu):
    import torch
    from abm_amd import _lib as L
    lib = L.load()
    rng = np.random.RandomState(3)
    R, n = 8200, 2
    ins = _inputs(rng, R, n, gpu)
    out = torch.zeros(6 * n, dtype=torch.float64, device=gpu)
    dphi = torch.zeros(n * R, dtype=torch.int8, device=gpu)
    rc = _call(lib, "abm_vf_state_variables_noalloc", R, n, *ins, (0.1, 0.09, 0.09), out, dphi)
    assert rc == L.ABM_ERR_INVALID
    assert b"8192" in lib.abm_last_error()
