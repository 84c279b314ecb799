"""The C-ABI library loads, exports every declared symbol, and its host-only
functions agree with the oracle (no GPU needed)."""
import os
import re

import numpy as np
import pytest

import oracle
from ecsprt_amd import _lib, configs
from ecsprt_amd import pic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ecsprt_hip.h")).read()
    return sorted(set(re.findall(r"\b(ecsprt_(?:hip|io)_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_struct_layouts_match_header():
    # Field order and count of the ctypes mirrors follow the header structs.
    text = open(os.path.join(ROOT, "include", "ecsprt_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for cname, mirror in (("ecsprt_params", _lib.Params), ("ecsprt_cli", _lib.CLI),
                          ("ecsprt_options", _lib.Options)):
        body = re.search(r"typedef struct \{([^{}]*?)\} " + cname + ";", text, re.S).group(1)
        fields = re.findall(r"\b(\w+)\s*[,;]", body)
        assert [f for f, _ in mirror._fields_] == fields, cname


def test_philox_known_answers():
    """Random123 known-answer vectors for Philox4x32-10."""
    assert pic.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert pic.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert pic.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox2x32_known_answers():
    """Random123 known-answer vectors for Philox2x32-10."""
    assert pic.philox2x32_10([0, 0], 0) == [0xFF1DAE59, 0x6CD10DF2]
    assert pic.philox2x32_10([0xFFFFFFFF] * 2, 0xFFFFFFFF) == [0x2C3F628B, 0xAB4FD7AD]
    assert pic.philox2x32_10([0x243F6A88, 0x85A308D3], 0x13198A2E) == [0xDD7CE038, 0xF62A4C12]


@pytest.mark.parametrize("wl", [configs.CPU32, configs.PROD128, configs.small(12)])
def test_setup_context_matches_oracle(wl):
    a = pic.setup_context(**wl.cli)
    b = oracle.setup_context(**wl.cli)
    assert a.ndim == 3
    for name, _ in a._fields_:
        if name != "ndim":
            assert getattr(a, name) == getattr(b, name), name


def test_setup_context_two_dimensional():
    """-ndim 2 (src/setup.c:178-188): Nz = 1 and dz from a unit z extent, every other
    parameter as in three dimensions; other ndim values are refused (src/setup.c:41)."""
    cli = dict(configs.small(12).cli, Nz=9, dz=0.3)
    a = pic.setup_context(**dict(cli, ndim=2))
    b = oracle.setup_context(**dict(cli, Nz=1, dz=1.0))
    assert a.ndim == 2 and a.Nz == 1 and a.dz == 1.0
    for name, _ in a._fields_:
        if name not in ("ndim", "z0", "z1", "Lz"):
            assert getattr(a, name) == getattr(b, name), name
    with pytest.raises(pic.HipError):
        pic.setup_context(**dict(cli, ndim=4))


@pytest.mark.parametrize("Tn_over_Ti,nui,np_", [(1.0, 0.02, 10 ** 6), (0.01, 0.01, 2097152), (1e-30, 0.01, 10 ** 6),
                                                 (100.0, 5.0, 1000)])
def test_collision_count_matches_oracle(Tn_over_Ti, nui, np_):
    cli = configs.base_cli(16, nui=nui, Tn_over_Ti=Tn_over_Ti)
    a = pic.setup_context(**cli)
    b = oracle.setup_context(**cli)
    assert pic.collision_count(a, np_, 1.0) == oracle.collision_count(b, np_, 1.0)


def test_workload_collision_counts():
    """The counts the workloads run with (recorded in the design notes)."""
    p0 = oracle.setup_context(**configs.CPU32.cli)
    p1 = oracle.setup_context(**configs.PROD128.cli)
    n0 = oracle.collision_count(p0, configs.CPU32.np, configs.CPU32.dt)
    n1 = oracle.collision_count(p1, configs.PROD128.np, configs.PROD128.dt)
    assert 0.3 * configs.CPU32.np < n0 < 0.5 * configs.CPU32.np
    assert 0.05 * configs.PROD128.np < n1 < 0.1 * configs.PROD128.np


# Index of the first ran1 output of at least 1 - 2^-24 (0.99999994f) from seed -12345.
# ran1 compares the float temp with 1 - DBL_EPSILON (src/random.c:10-11,67-68), so this
# value passes through unchanged; the Numerical Recipes book's 1 - 1.2e-7 would return
# 0.99999988f here.
RAN1_FIRST_NEAR_ONE = 16392033


@pytest.mark.parametrize("gen,n", [("ran1", RAN1_FIRST_NEAR_ONE + 200_000), ("ran3", 3_000_000),
                                   ("gasdev", 14_000_000)])
def test_library_nr_streams_equal_oracle(gen, n):
    """The library's host ran1 / ran3 / gasdev streams (the reference's initial conditions
    are drawn from them) equal the oracle's value for value, through the first ran1 output
    that rounds to 1 - 2^-24 (gasdev's 14M deviates consume about 17.8M ran1 values)."""
    ours = pic.nr_stream(gen, -12345, n)
    ref = oracle.nr_stream(gen, -12345, n)
    np.testing.assert_array_equal(ours, ref)
    if gen == "ran1":
        near_one = np.nonzero(ref >= np.float32(1) - np.float32(2.0 ** -24))[0]
        assert near_one[0] == RAN1_FIRST_NEAR_ONE
        assert ref[RAN1_FIRST_NEAR_ONE] == np.float32(1) - np.float32(2.0 ** -24)


def test_nr_stream_rejects_bad_generator():
    import ctypes
    out = np.zeros(1, np.float32)
    rc = _lib.load().ecsprt_hip_nr_stream(3, -1, 1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert rc == _lib.ERR_ARG


@pytest.mark.parametrize("lam", [0.0, 1e-3, 0.08, 1.0, 2.0, 16.5, 63.7, 100.0, 5000.0, 2.5e5])
def test_poisson_attempts_equal_quantile(lam):
    """The Philox mode's attempt count is the Poisson(lam) quantile of its uniform, value
    for value against scipy's, from a tenth of an attempt per ion (the 128^3 workload) to
    the rates a cooling cold-neutral run reaches (16.5 per ion at 32^3 step 10) and far
    beyond: no truncation of the tail at any rate."""
    from scipy import stats
    rng = np.random.default_rng(1)
    u = (rng.integers(0, 2 ** 24, 200_000) + 0.5) / 2 ** 24  # the kernels' 24-bit uniforms
    u = np.concatenate([u, [0.5 / 2 ** 24, 1 - 0.5 / 2 ** 24]])  # the extreme draws
    k = pic.poisson_attempts(lam, u)
    ref = stats.poisson.ppf(u, lam).astype(np.int64) if lam > 0 else np.zeros(u.size, np.int64)
    if lam <= 5000.0:
        np.testing.assert_array_equal(k, ref)
    else:
        # hundreds of FP64 CDF terms summed from the mode: a draw within rounding of a CDF
        # value may land one count over, in either implementation
        assert np.abs(k - ref).max() <= 1 and np.count_nonzero(k != ref) <= 5


def test_poisson_attempts_rate_bound():
    with pytest.raises(pic.HipError) as e:
        pic.poisson_attempts(2e6, np.array([0.5]))
    assert e.value.code == _lib.ERR_COLLISION
