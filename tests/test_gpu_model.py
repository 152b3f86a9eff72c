"""The "hip" plug-in model through the reference's ModelFunctions boundary
(src/model_interface.h:30-36), called the way src/main_runner.c:198-298 calls a
model: GSL-layout params owned by the caller, zero-initialised result,
S_values freed by the caller, -1 on failure."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from helpers import factor_like

pytestmark = pytest.mark.gpu
libc = ctypes.CDLL(None)
libc.free.argtypes = [ctypes.c_void_p]


def test_get_hip_model_record(mcr):
    m = mcr.get_hip_model()
    assert m.name == b"hip"


@pytest.mark.parametrize("stride,tda", [(1, None), (3, 17)])
def test_run_model_matches_c_abi(mcr, stride, tda):
    mu, sigma = factor_like(12, 3, 61)
    p, keep = mcr.make_params(mu, sigma, 0.02, 3, 50_000, 67, mu_stride=stride, sigma_tda=tda)
    res = mcr.MonteCarloResult()
    m = mcr.get_hip_model()
    assert m.run_model(ctypes.byref(p), ctypes.byref(res)) == 0
    ref = mcr.run(mu, sigma, 0.02, 3, 50_000, 67)
    assert res.count == ref.count
    assert res.P_hat == ref.p_hat and res.std_error == ref.std_error
    assert res.ci_lower == ref.ci_lower and res.ci_upper == ref.ci_upper
    assert not res.S_values  # NULL unless requested
    # caller-owned inputs untouched
    assert np.array_equal(keep[0][::stride], mu)


def test_run_model_fills_s_values_on_request(mcr, monkeypatch):
    mu, sigma = factor_like(9, 2, 62)
    monkeypatch.setenv("MCR_HIP_S_VALUES", "1")
    p, keep = mcr.make_params(mu, sigma, 0.02, 2, 4096, 5)
    res = mcr.MonteCarloResult()
    assert mcr.lib().hip_simulate(ctypes.byref(p), ctypes.byref(res)) == 0
    S = np.ctypeslib.as_array(res.S_values, shape=(4096,)).copy()
    libc.free(ctypes.cast(res.S_values, ctypes.c_void_p))
    ref = mcr.run(mu, sigma, 0.02, 2, 4096, 5, want_S=True)
    assert np.array_equal(S, ref.S)
    assert res.count == int(np.sum(S >= 2))


def test_run_model_errors_return_minus_one(mcr):
    mu, sigma = factor_like(6, 2, 63)
    p, keep = mcr.make_params(mu, sigma, 0.02, 2, 1000, 5)
    p.N = 7  # does not match mu / Sigma
    res = mcr.MonteCarloResult()
    assert mcr.lib().hip_simulate(ctypes.byref(p), ctypes.byref(res)) == -1
    p.N = 6
    p.M = 0
    assert mcr.lib().hip_simulate(ctypes.byref(p), ctypes.byref(res)) == -1


def test_c_abi_argument_errors(mcr):
    mu, sigma = factor_like(4, 2, 64)
    with pytest.raises(mcr.McrError) as e:
        mcr.run(mu, sigma, 0.02, 1, 0, 1)
    assert e.value.code == mcr.MCR_E_ARG
    with pytest.raises(mcr.McrError) as e:
        mcr.run(mu, sigma, 0.02, 1, 10, 1, kernel=mcr.KERNEL_MFMA, alphas=(1.5,))
    assert e.value.code == mcr.MCR_E_ARG
    big = np.eye(1025) * 1e-4
    with pytest.raises(mcr.McrError) as e:
        mcr.run(np.zeros(1025), big, 0.02, 1, 10, 1)
    assert e.value.code == mcr.MCR_E_UNSUPPORTED
    with pytest.raises(mcr.McrError) as e:
        mcr.run(np.zeros(100), np.eye(100) * 1e-4, 0.02, 1, 10, 1, kernel=mcr.KERNEL_VALU_REG)
    assert e.value.code == mcr.MCR_E_UNSUPPORTED
    nan_sigma = sigma.copy()
    nan_sigma[0, 1] = np.nan
    with pytest.raises(mcr.McrError) as e:
        mcr.run(mu, nan_sigma, 0.02, 1, 10, 1)
    assert e.value.code == mcr.MCR_E_FACTOR


def test_reference_python_api_mirror(mcr):
    mu, sigma = factor_like(10, 2, 65)
    P_hat, count, S = mcr.monte_carlo_crash_simulation(0.02, 10, 3, mu, sigma, 10000, seed=9)
    assert count == int(np.sum(S >= 3)) and P_hat == count / 10000
    se, lo, hi = mcr.compute_accuracy(P_hat, 10000)
    assert lo <= P_hat <= hi and se > 0
