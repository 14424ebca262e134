"""binary16 conversion and the inflow hash of the oracle."""
import numpy as np


def test_half_roundtrip_exhaustive(oracle_lib):
    lib = oracle_lib
    bad = 0
    for h in range(0, 65536, 1):
        e = (h >> 10) & 0x1F
        if e == 31 and (h & 0x3FF):
            continue  # NaN payloads
        f = lib.pfo_h2f(h)
        if lib.pfo_f2h(f) != h:
            bad += 1
    assert bad == 0


def test_half_matches_numpy_rne(oracle_lib):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=20000) * 10.0 ** rng.integers(-9, 5, 20000),
                        np.array([65504.0, 65519.99, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8, -0.0])])
    x = x.astype(np.float32)
    want = x.astype(np.float16).view(np.uint16)
    got = np.array([oracle_lib.pfo_f2h(float(v)) for v in x], dtype=np.uint16)
    np.testing.assert_array_equal(got, want)


def _lowbias32(x):
    # Published lowbias32 integer hash (C. Wellons): xorshift-multiply, twice.
    x = np.uint64(x)
    m = np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & m
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & m
    x ^= x >> np.uint64(16)
    return int(x)


def test_hash32(oracle_lib):
    for v in [0, 1, 2, 0x9E3779B9, 0xFFFFFFFF, 123456789, 0xA341316C ^ 0x9E3779B9]:
        assert oracle_lib.pfo_hash32(v) == _lowbias32(v)
    assert oracle_lib.pfo_hash32(0) == 0
