"""Deterministic-mode arithmetic of include/bg_detmath.h (shared by the kernels and the oracle).

Philox4x32-10 is pinned by the known-answer vectors published with Random123 (Salmon et al.,
SC'11); the double-precision functions by comparison with the C library at <= 2 ulp; bg_expf at
<= 1 ulp of the correctly rounded value; the noise transform by its moments and a
Kolmogorov-Smirnov test against N(0,1).
"""
import ctypes

import numpy as np
import pytest
from scipy import stats

from oracle import oracle_py as O

L = O.lib()
L.bgo_philox.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
L.bgo_noise.restype = ctypes.c_float
L.bgo_noise.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
L.bgo_detmath.restype = ctypes.c_double
L.bgo_detmath.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]

KAT = [  # counter, key, expected output (Random123 kat_vectors, philox4x32 R=10)
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    L.bgo_philox(c, k, o)
    assert tuple(o) == want


def _ulp_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    sp = np.spacing(np.abs(b))
    return np.abs(a - b) / np.where(sp > 0, sp, 5e-324)


@pytest.mark.parametrize("fn,lo,hi,ref,tol", [
    (0, -700.0, 700.0, np.exp, 2), (1, 1e-300, 1e300, np.log, 2), (4, -1.0, 1.0, np.arccos, 2),
])
def test_double_functions_within_two_ulp(fn, lo, hi, ref, tol):
    rng = np.random.default_rng(fn)
    if fn == 1:
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), 20000))
    else:
        x = rng.uniform(lo, hi, 20000)
    got = np.array([L.bgo_detmath(fn, float(v), 0.0) for v in x])
    assert _ulp_err(got, ref(x)).max() <= tol


def test_sin_cos_absolute_error():
    x = np.linspace(-7, 7, 20001)
    s = np.array([L.bgo_detmath(2, float(v), 0.0) for v in x])
    c = np.array([L.bgo_detmath(3, float(v), 0.0) for v in x])
    assert np.abs(s - np.sin(x)).max() < 4e-16
    assert np.abs(c - np.cos(x)).max() < 4e-16
    u = np.linspace(0, 1, 20001, endpoint=False)
    c2 = np.array([L.bgo_detmath(6, float(v), 0.0) for v in u])
    assert np.abs(c2 - np.cos(2 * np.pi * u)).max() < 1e-15


def test_acos_special_values():
    assert L.bgo_detmath(4, 1.0, 0) == 0.0
    assert L.bgo_detmath(4, -1.0, 0) == np.pi
    assert np.isnan(L.bgo_detmath(4, 1.0000001, 0))


def test_pow_special_cases():
    assert L.bgo_detmath(5, 0.0, 0.0) == 1.0
    assert L.bgo_detmath(5, -3.0, 0.0) == 1.0
    assert L.bgo_detmath(5, 0.0, 2.0) == 0.0
    assert abs(L.bgo_detmath(5, 2.0, 10.0) - 1024.0) < 1e-12
    assert abs(L.bgo_detmath(5, -2.0, 3.0) + 8.0) < 1e-12


def test_expf_one_ulp():
    x = np.random.default_rng(3).uniform(-103.0, 88.7, 50000).astype(np.float32)
    got = np.array([L.bgo_detmath(7, float(v), 0.0) for v in x], np.float32)
    want = np.exp(x.astype(np.float64)).astype(np.float32)
    normal = want > np.float32(1.2e-38)
    ulp32 = np.abs(got[normal].astype(np.float64) - want[normal]) / np.spacing(want[normal]).astype(np.float64)
    assert ulp32.max() <= 1.0
    assert L.bgo_detmath(7, 0.0, 0) == 1.0


def test_noise_is_standard_normal():
    z = np.array([L.bgo_noise(1, i % 1000, i // 1000) for i in range(200000)], np.float64)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3


def test_noise_is_a_pure_function_of_seed_neuron_step():
    a = [L.bgo_noise(7, 5, 123456789) for _ in range(3)]
    assert a[0] == a[1] == a[2]
    assert L.bgo_noise(7, 5, 1) != L.bgo_noise(7, 6, 1)
    assert L.bgo_noise(7, 5, 1) != L.bgo_noise(8, 5, 1)
    assert L.bgo_noise(7, 5, 1) != L.bgo_noise(7, 5, 2)
