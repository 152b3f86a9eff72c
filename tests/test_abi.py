"""C-ABI surface (CPU): the built library loads, exports every function that
include/mcr_hip.h declares, and the host-only data generator behaves.  No GPU
call is made here."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest


def test_library_exports_every_declared_function(mcr):
    names = mcr.declared_functions()
    assert {"mcr_run", "mcr_factor", "get_hip_model", "hip_simulate", "mcr_server_submit",
            "mcr_read_mu_bin", "mcr_load_config", "mcr_append_csv"} <= set(names)
    L = ctypes.CDLL(mcr.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_model_interface_header_layout(mcr):
    # field offsets of the GSL-layout structs the reference driver passes in
    assert ctypes.sizeof(mcr.GslVector) == 40
    assert mcr.GslMatrix.data.offset == 24
    assert mcr.MonteCarloParams.M.offset == 32
    assert mcr.MonteCarloResult.S_values.offset == 16


def test_synth_matches_numpy_moments(mcr):
    mu, sigma, R = mcr.synth_factor_model(12, 3, 700, 77, want_returns=True)
    assert np.allclose(mu, R.mean(axis=0), rtol=0, atol=1e-15)
    assert np.allclose(sigma, np.cov(R, rowvar=False, ddof=1), rtol=1e-12, atol=1e-18)
    assert np.array_equal(sigma, sigma.T)
    assert np.all(np.linalg.eigvalsh(sigma) > 0)


def test_synth_is_deterministic_and_seeded(mcr):
    a = mcr.synth_factor_model(8, 2, 300, 5)
    b = mcr.synth_factor_model(8, 2, 300, 5)
    c = mcr.synth_factor_model(8, 2, 300, 6)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[1], c[1])


def test_synth_factor_structure(mcr):
    # population covariance is B B^T fvol^2 + ivol^2 I with B ~ U(0.2, 0.8):
    # mean diagonal ~ K E[B^2] 1e-4 + 1e-4
    n, K, T = 40, 5, 20000
    mu, sigma = mcr.synth_factor_model(n, K, T, 1002)
    eb2 = (0.8**3 - 0.2**3) / (3 * 0.6)
    expect_diag = K * eb2 * 1e-4 + 1e-4
    assert np.mean(np.diag(sigma)) == pytest.approx(expect_diag, rel=0.15)
    assert np.mean(mu) == pytest.approx(2e-4, abs=3 * np.sqrt(np.mean(np.diag(sigma)) / T))


@pytest.mark.skipif(os.environ.get("MCR_SLOW") != "1", reason="n=600 generator (set MCR_SLOW=1)")
def test_synth_large_condition_number(mcr):
    mu, sigma = mcr.synth_factor_model(600, 20, 2500, 1003)
    c = np.linalg.cond(sigma)
    assert 1e3 < c < 1e6


def test_bad_arguments_rejected_without_gpu(mcr):
    with pytest.raises(mcr.McrError):
        mcr.synth_factor_model(0, 2, 100, 1)


def test_misshaped_arrays_rejected_before_the_library(mcr):
    """The C side copies n rows of ld_sigma doubles and n weights from the
    caller's pointers; arrays too small for that must fail in Python."""
    mu, sigma = np.zeros(6), np.eye(6)
    ok, _ = mcr._make_query(mu, sigma, 0.02, 2, 64, 1)
    assert ok.n == 6 and ok.ld_sigma == 6
    strided, _ = mcr._make_query(mu, np.eye(6, 9), 0.02, 2, 64, 1)
    assert strided.ld_sigma == 9
    for bad in (dict(sigma=np.eye(5)), dict(sigma=np.eye(6)[:5]), dict(sigma=np.zeros(35)),
                dict(weights=np.ones(5) / 5), dict(weights=np.ones((6, 1))),
                dict(mu=np.zeros((6, 1)))):
        args = dict(mu=mu, sigma=sigma, weights=None)
        args.update(bad)
        with pytest.raises(ValueError):
            mcr._make_query(args["mu"], args["sigma"], 0.02, 2, 64, 1, weights=args["weights"])
    with pytest.raises(ValueError):
        mcr.factor(np.eye(6)[:5])
    with pytest.raises(ValueError):
        mcr.prepare(mu, np.eye(5), 0.02)


def test_server_footprint_and_servable_checks(mcr):
    """Host-only admission arithmetic of the serving loop."""
    n = 600
    p = {"mu": np.zeros(n), "sigma": np.eye(n), "x": 0.02, "k": 5, "paths": 10**7, "seed": 1,
         "alphas": (0.99,), "precision": mcr.PREC_BF16X3}
    b = mcr.server_footprint(p)
    # factor (FP64 L) + covariance copy + bf16 hi/lo operands + one VaR buffer
    # dominate; the VaR buffer holds the window's expected count (the loss is
    # Gaussian: window za +- ha standard deviations) plus 8 sigma and 64
    from scipy.stats import norm
    T, S = (n + 1 + 31) // 32, (n + 15) // 16
    za = norm.ppf(0.99)
    ha = 6.0 * np.sqrt(0.99 * 0.01 / 1e7) / norm.pdf(za) + 1e-4
    eg = 1e7 * (norm.cdf(za + ha) - norm.cdf(za - ha))
    wcap = int(np.ceil(eg + 8 * np.sqrt(eg) + 64))
    assert 3000 < eg < 5000 and wcap < 65536
    base = 2 * n * n * 8 + T * S * 128 * 16
    assert b >= base + wcap * 4
    assert b <= 1.1 * (base + wcap * 4) + 8192
    # windows that hold every loss get the whole 2^16-loss buffer
    wide = mcr.server_footprint(dict(p, extra_flags=mcr.F_TEST_WIDE))
    assert wide >= base + 65536 * 4 and wide - b >= (65536 - wcap) * 4 - 256
    small = dict(p, mu=np.zeros(5), sigma=np.eye(5), alphas=())
    assert mcr.server_footprint(small) < 16384
    # served queries take up to MCR_MAX_N_BATCH assets (24 row tiles)
    widest = dict(p, mu=np.zeros(767), sigma=np.eye(767), weights=None)
    assert mcr.server_footprint(widest) > 2 * 767 * 767 * 8
    over = dict(p, mu=np.zeros(768), sigma=np.eye(768), weights=None)
    assert mcr.MCR_MAX_N_BATCH == 767
    for bad, code in (({"precision": mcr.PREC_FP32}, mcr.MCR_E_UNSUPPORTED),
                      ({"extra_flags": mcr.F_NEAR}, mcr.MCR_E_UNSUPPORTED),
                      ({"paths": 0}, mcr.MCR_E_ARG),
                      ({"mu": over["mu"], "sigma": over["sigma"], "weights": None},
                       mcr.MCR_E_UNSUPPORTED)):
        with pytest.raises(mcr.McrError) as e:
            mcr.server_footprint(dict(p, **bad))
        assert e.value.code == code
    mu = np.zeros(n)
    mu[7] = np.inf
    with pytest.raises(mcr.McrError) as e:
        mcr.server_footprint(dict(p, mu=mu))
    assert e.value.code == mcr.MCR_E_ARG


def test_noalloc_entry_point_needs_init(mcr):
    """mcr_run_noalloc creates nothing itself: before mcr_init (here no GPU is
    ever initialised) it refuses the query without touching a device."""
    q, keep = mcr._make_query(np.zeros(4), np.eye(4) * 1e-4, 0.02, 2, 64, 1)
    r = mcr.Result()
    code = mcr.lib().mcr_run_noalloc(ctypes.byref(q), ctypes.byref(r))
    assert code == mcr.MCR_E_ARG
    assert b"mcr_init" in mcr.lib().mcr_last_error()
    del keep


def test_ctypes_structs_match_the_c_header(mcr, tmp_path):
    """Every struct of include/mcr_hip.h that the Python view mirrors has the
    same size and field offsets as gcc lays it out (a C probe compiled here)."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    pairs = {"mcr_query": mcr.Query, "mcr_result": mcr.Result, "mcr_factor_info": mcr.FactorInfo,
             "mcr_model_export": mcr.ModelExport, "mcr_batch_opts": mcr.BatchOpts,
             "mcr_server_opts": mcr.ServerOpts, "mcr_ticket_info": mcr.TicketInfo,
             "mcr_server_stats": mcr.ServerStats}
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mcr_hip.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            cname, what, val = line.split()
            got[(cname, what)] = int(val)
    for cname, py in pairs.items():
        assert ctypes.sizeof(py) == got[(cname, "size")], cname
        for fname, _ in py._fields_:
            assert getattr(py, fname).offset == got[(cname, fname)], (cname, fname)


def test_python_constants_match_the_c_header(mcr):
    """Every MCR_* integer constant the Python view defines (limits, error
    codes, precisions, kernel ids, flags) equals the header's #define."""
    import re

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "mcr_hip.h")).read()
    defs = {m.group(1): int(m.group(2), 0)
            for m in re.finditer(r"^#define (MCR_\w+) (-?(?:0x[0-9a-fA-F]+|\d+))\b", hdr, re.M)}
    checked = 0
    for name, value in defs.items():
        py = name if hasattr(mcr, name) else name[len("MCR_"):]
        if hasattr(mcr, py):
            assert getattr(mcr, py) == value, (name, getattr(mcr, py), value)
            checked += 1
    for name in ("MCR_MAX_N", "MCR_MAX_N_BATCH", "MCR_MAX_N_NOALLOC", "MCR_KERNEL_VALU_CHUNK",
                 "MCR_E_UNSUPPORTED", "MCR_PREC_AUTO", "MCR_F_NEAR"):
        assert name in defs
    assert checked >= 15
