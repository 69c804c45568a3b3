"""The C ABI library: loads, exports every declared symbol, validates input
on the host (no GPU needed for any of these)."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from abm_amd import _lib as L


def test_library_loads_and_reports_build():
    lib = L.load()
    info = lib.abm_build_info().decode()
    assert "gfx950" in info


def test_every_declared_symbol_is_exported():
    declared = L.declared_symbols()
    assert len(declared) >= 19
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    # and the ctypes signature table covers them all
    assert set(declared) <= set(L._SIGS)


def _rep(**kw):
    base = dict(n_agents=4, n_rays=64, boundary=0, limit_movement=0, fov0=-np.pi, fov1=np.pi,
                width=100.0, height=100.0, window_pad=30.0, gam=0.1, v0=1.0, alp0=1.0, alp1=0.09,
                bet0=1.0, bet1=0.09, max_vel=3.0, max_th=0.1)
    base.update(kw)
    return L.VFReplica(**base)


@pytest.mark.parametrize("kw, needle", [
    (dict(n_rays=1), "n_rays"), (dict(n_agents=0), "n_agents"), (dict(boundary=7), "boundary"),
    (dict(alp0=float("nan")), "non-finite"), (dict(width=0.0), "arena size")])
def test_batch_create_rejects_bad_replicas(kw, needle):
    lib = L.load()
    arr = (L.VFReplica * 2)(_rep(), _rep(**kw))
    h = C.c_void_p()
    rc = lib.abm_vf_batch_create(arr, 2, C.byref(h))
    assert rc == L.ABM_ERR_INVALID
    assert needle in lib.abm_last_error().decode()
    assert not h.value


def test_step_rejects_null_arguments():
    lib = L.load()
    rc = lib.abm_vf_step_f64(None, None, None, None, 0, None)
    assert rc == L.ABM_ERR_INVALID
    rc = lib.abm_vf_projection_rows(1, 0.0, 0.0, 0, 1.0, 1.0, 0, 0, 0, 1, None, None, None, 0,
                                    None, None, None)
    assert rc == L.ABM_ERR_INVALID


def test_struct_layouts():
    assert C.sizeof(L.VFReplica) == 4 * 4 + 13 * 8
    assert C.sizeof(L.BatchInfo) == 24
    assert C.sizeof(L.VFStateF64) == 8 * 8


def test_bit_packing_roundtrip():
    from abm_amd.vf import pack_bits, unpack_bits
    rng = np.random.RandomState(0)
    for R in (8, 31, 32, 33, 512, 1000):
        b = (rng.rand(3, R) < 0.3).astype(np.uint8)
        assert np.array_equal(unpack_bits(pack_bits(b), R), b)
