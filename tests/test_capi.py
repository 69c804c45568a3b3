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


def _addrs(gens):
    return np.array([g._bit_generator.ctypes.state_address for g in gens], dtype=np.uint64)


def test_mt19937_peek_and_skip_follow_numpys_legacy_stream():
    # host-only helpers of the random-walk protocol: peek must equal draws
    # from a copy of each RandomState, skip must leave each generator where
    # random_sample(count) leaves it; positions straddle the 624-word twist
    lib = L.load()
    gens = [np.random.RandomState(s) for s in range(40)]
    for i, g in enumerate(gens):
        g.random_sample((i * 37) % 700)
    gens[0].get_state()                       # fresh twist pending (pos 624) on some
    counts = np.array([(i * 53) % 1500 for i in range(40)], dtype=np.int64)
    addr = _addrs(gens)
    out = np.zeros(int(counts.sum()))
    before = [g.get_state() for g in gens]
    assert lib.abm_mt19937_peek(addr.ctypes.data_as(C.c_void_p), len(gens),
                                counts.ctypes.data_as(C.c_void_p),
                                out.ctypes.data_as(C.c_void_p)) == 0
    want = []
    for g, c in zip(gens, counts):
        t = np.random.RandomState()
        t.set_state(g.get_state())
        want.append(t.random_sample(int(c)))
    assert np.array_equal(out, np.concatenate(want))
    for g, b in zip(gens, before):               # peek changed nothing
        assert g.get_state()[2] == b[2] and np.array_equal(g.get_state()[1], b[1])
    assert lib.abm_mt19937_skip(addr.ctypes.data_as(C.c_void_p), len(gens),
                                counts.ctypes.data_as(C.c_void_p)) == 0
    for g, b, c in zip(gens, before, counts):
        t = np.random.RandomState()
        t.set_state(b)
        t.random_sample(int(c))
        assert g.get_state()[2] == t.get_state()[2]
        assert np.array_equal(g.get_state()[1], t.get_state()[1])
        assert g.uniform(-0.3, 0.3) == t.uniform(-0.3, 0.3)


def test_mt19937_peek32_and_skip32_give_numpys_raw_words():
    # the resource loop's words: peek32 must equal next_uint32 of a copy of
    # each RandomState (randint over the full 32-bit range returns them
    # unmasked), skip32 must leave it where that many words leave it
    lib = L.load()
    gens = [np.random.RandomState(100 + s) for s in range(12)]
    for i, g in enumerate(gens):
        g.random_sample((i * 71) % 400)
    counts = np.array([(i * 97) % 1400 for i in range(12)], dtype=np.int64)
    addr = _addrs(gens)
    out = np.zeros(int(counts.sum()), dtype=np.uint32)
    before = [g.get_state() for g in gens]
    assert lib.abm_mt19937_peek32(addr.ctypes.data_as(C.c_void_p), len(gens),
                                  counts.ctypes.data_as(C.c_void_p),
                                  out.ctypes.data_as(C.c_void_p)) == 0
    want = []
    for g, c in zip(gens, counts):
        t = np.random.RandomState()
        t.set_state(g.get_state())
        want.append(t.randint(0, 2 ** 32, size=int(c), dtype=np.uint32))
    assert np.array_equal(out, np.concatenate(want))
    assert lib.abm_mt19937_skip32(addr.ctypes.data_as(C.c_void_p), len(gens),
                                  counts.ctypes.data_as(C.c_void_p)) == 0
    for g, b, c in zip(gens, before, counts):
        t = np.random.RandomState()
        t.set_state(b)
        t.randint(0, 2 ** 32, size=int(c), dtype=np.uint32)
        assert g.get_state()[2] == t.get_state()[2]
        assert np.array_equal(g.get_state()[1], t.get_state()[1])


def test_world_struct_layouts():
    assert C.sizeof(L.ForageWorldParams) == 10 * 4 + 3 * 8
    assert L.ForageWorldParams.consumption.offset == 40
    assert C.sizeof(L.PatchState) == 8 * 8


def test_mt19937_rejects_bad_arguments():
    lib = L.load()
    g = np.random.RandomState(1)
    addr = _addrs([g])
    neg = np.array([-1], dtype=np.int64)
    assert lib.abm_mt19937_skip(addr.ctypes.data_as(C.c_void_p), 1,
                                neg.ctypes.data_as(C.c_void_p)) == L.ABM_ERR_INVALID
    assert b"negative" in lib.abm_last_error()
    one = np.array([1], dtype=np.int64)
    assert lib.abm_mt19937_peek(addr.ctypes.data_as(C.c_void_p), 1,
                                one.ctypes.data_as(C.c_void_p), None) == L.ABM_ERR_INVALID
