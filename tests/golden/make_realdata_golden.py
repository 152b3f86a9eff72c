"""Writes realdata_stats.json: the preprocessing outputs for the reference's
own index data (dataset/*.csv, 15 European indices) under its config.yaml,
computed here with pandas through the same library calls the reference
pipeline makes (read_csv; to_datetime with format %m/%d/%Y and errors=coerce;
a DataFrame of the Return series aligned on their dates; .loc[start:end];
dropna; int(len * train_ratio) split; mean; cov; (test < -x).sum(axis=1) >= k).
Pins the numpy restatement in gpu-montecarlo-risk_amd/preprocess.py.

Run from the repository root where the reference tree is present:
    python tests/golden/make_realdata_golden.py [/root/reference]
"""
import json
import os
import sys

import numpy as np
import pandas as pd
import yaml

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
with open(os.path.join(REF, "config.yaml")) as f:
    cfg = yaml.safe_load(f)

series = {}
for name, fname in cfg["indices"].items():
    df = pd.read_csv(os.path.join(REF, "dataset", fname))
    df["Date"] = pd.to_datetime(df["Date"], format="%m/%d/%Y", errors="coerce")
    df = df.dropna(subset=["Date"]).set_index("Date")
    series[name] = df["Return"]
returns = pd.DataFrame(series)
returns = returns.loc[pd.to_datetime(cfg["start"]):pd.to_datetime(cfg["end"])].dropna()
split = int(len(returns) * cfg["train_ratio"])
train, test = returns.iloc[:split], returns.iloc[split:]
mu = train.mean().values
sigma = train.cov().values
freq = float(np.mean((test < -cfg["x"]).sum(axis=1) >= cfg["k"]))

out = {"source": "dataset/*.csv and config.yaml of the reference tree",
       "config": {"start": cfg["start"], "end": cfg["end"], "x": cfg["x"], "k": cfg["k"],
                  "train_ratio": cfg["train_ratio"], "M": cfg["M"],
                  "random_seed": cfg["random_seed"], "indices": cfg["indices"]},
       "names": list(returns.columns), "days": len(returns), "train_days": len(train),
       "test_days": len(test), "first_date": str(returns.index[0].date()),
       "last_date": str(returns.index[-1].date()), "actual_freq": freq,
       "mu": [float(v) for v in mu], "Sigma": [[float(v) for v in row] for row in sigma],
       "cond_sigma": float(np.linalg.cond(sigma))}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "realdata_stats.json")
with open(path, "w") as f:
    json.dump(out, f, indent=1)
print(f"wrote {path}: {out['days']} days, {len(out['names'])} indices, "
      f"actual_freq {freq}, cond {out['cond_sigma']:.4g}")
