"""Real-data preprocessing: index return CSVs -> (mu, Sigma, actual crash
frequency) -> mu.bin / sigma.bin / params.bin for the risk path.

Restates the reference pipeline (plain Python + numpy, no pandas):
  run_pipeline      preprocessing/main.py:8-50
  fetch_and_split   preprocessing/fetch_split.py:6-124
  compute_stats     preprocessing/preprocessing_stats.py:5-28 (mean, ddof=1 covariance)
  compute_metrics   preprocessing/evaluation.py:5-31
  write_bins        preprocessing/save_binary.py:5-45, through mcr_write_bins (C ABI)

Semantics kept: dates parsed strictly as %m/%d/%Y (rows whose date does not
parse are dropped), indices aligned on the union of their dates in ascending
order, the [start, end] slice inclusive at both ends, rows with any missing
return dropped, at least 100 rows, split index int(rows * train_ratio), crash
iff return < -x, an extreme day iff at least k crashes.
"""
from __future__ import annotations

import csv
import datetime as _dt
import math
import os

import numpy as np


class PreprocessError(ValueError):
    pass


def _parse_date(s: str):
    s = s.strip()
    parts = s.split("/")
    if len(parts) != 3 or not all(p.isdigit() for p in parts) or len(parts[2]) != 4:
        return None
    try:
        return _dt.date(int(parts[2]), int(parts[0]), int(parts[1]))
    except ValueError:
        return None


def _parse_float(s: str) -> float:
    s = s.strip()
    if s == "":
        return math.nan
    try:
        return float(s)
    except ValueError:
        return math.nan


def load_index_csv(path: str) -> dict | None:
    """{date: return} of one index file, or None when the file lacks the
    Date/Return columns.  Rows whose date does not parse are dropped."""
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    if not rows:
        return None
    header = [h.strip() for h in rows[0]]
    if "Date" not in header or "Return" not in header:
        return None
    di, ri = header.index("Date"), header.index("Return")
    out = {}
    for r in rows[1:]:
        if len(r) <= max(di, ri):
            continue
        d = _parse_date(r[di])
        if d is None:
            continue
        if d in out:
            raise PreprocessError(f"{path}: duplicate date {d}")
        out[d] = _parse_float(r[ri])
    return out


def _to_date(v) -> _dt.date:
    if isinstance(v, _dt.datetime):
        return v.date()
    if isinstance(v, _dt.date):
        return v
    return _dt.date.fromisoformat(str(v).strip()[:10])


def fetch_and_split(indices: dict, dataset_dir: str, start="2015-01-01", end="2025-01-01",
                    train_ratio=0.7, log=print):
    """(train, test, dates, names): train/test are [days, indices] float64."""
    series, names = [], []
    for name, fname in indices.items():
        path = os.path.join(dataset_dir, fname)
        if not os.path.exists(path):
            log(f"  WARNING: CSV file not found: {fname} (skipping {name})")
            continue
        s = load_index_csv(path)
        if s is None:
            log(f"  WARNING: {fname} missing 'Date' or 'Return' column (skipping {name})")
            continue
        series.append(s)
        names.append(name)
    if not series:
        raise PreprocessError("no index data could be loaded")
    dates = sorted(set().union(*[s.keys() for s in series]))
    lo, hi = _to_date(start), _to_date(end)
    dates = [d for d in dates if lo <= d <= hi]
    if not dates:
        raise PreprocessError(f"no data in the range {start} to {end}")
    R = np.array([[s.get(d, math.nan) for s in series] for d in dates], dtype=np.float64)
    keep = ~np.isnan(R).any(axis=1)
    R = R[keep]
    dates = [d for d, k in zip(dates, keep) if k]
    if len(R) < 100:
        raise PreprocessError(f"only {len(R)} days after alignment (minimum 100)")
    if not 0.0 < train_ratio < 1.0:
        raise PreprocessError(f"train_ratio={train_ratio} must be in (0, 1)")
    split = int(len(R) * train_ratio)
    return R[:split], R[split:], dates, names


def compute_stats(train: np.ndarray):
    """Sample mean and ddof=1 covariance of the training returns."""
    return train.mean(axis=0), np.cov(train, rowvar=False, ddof=1)


def compute_metrics(test: np.ndarray, x=0.02, k=5) -> float:
    """Fraction of test days with at least k indices below -x."""
    crashes = (test < -x).sum(axis=1)
    return float(np.mean(crashes >= k))


def write_bins(out_dir: str, mu, sigma, actual_freq: float) -> None:
    """mu.bin / sigma.bin / params.bin through the C ABI (mcr_write_bins)."""
    from . import write_bins as _write

    _write(out_dir, mu, sigma, actual_freq)


def run_pipeline(config: dict, dataset_dir: str, out_dir: str | None = None, log=print) -> dict:
    """config: the parsed config.yaml (start, end, x, k, train_ratio, indices)."""
    indices = config["indices"]
    if not isinstance(indices, dict):
        raise TypeError("'indices' must map index names to CSV file names")
    train, test, dates, names = fetch_and_split(indices, dataset_dir, config["start"],
                                                config["end"], config["train_ratio"], log)
    mu, sigma = compute_stats(train)
    actual_freq = compute_metrics(test, config["x"], config["k"])
    if out_dir is not None:
        write_bins(out_dir, mu, sigma, actual_freq)
    return {"mu": mu, "Sigma": sigma, "actual_freq": actual_freq, "names": names,
            "train_days": len(train), "test_days": len(test), "first_date": dates[0],
            "last_date": dates[-1]}
