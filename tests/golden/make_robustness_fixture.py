"""Writes tests/golden/robustness_runs.npz from the reference's FP32
robustness protocol data (notebooks/robustness_data/ of the reference, read
as data with numpy's non-pickling loader):

* ys (100, 7500, 2) float32: data-T8000-Run{1..100}.npz, the 100 simulated
  bearings-only runs, truncated to the protocol's longest prefix (7500);
* Ts (24,): the prefix lengths of notebooks/robustness_100runs.py:48;
* ell_{extended,cubature}_{sqrt,std} (24, 100) float32: the reference's FP32
  log-likelihoods after 20 iterations (ell_float32_*_100runs.npz), NaN where
  the run failed.

Run once in a container that holds the reference: python tests/golden/make_robustness_fixture.py
"""
import os

import numpy as np

SRC = "/root/reference/notebooks/robustness_data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "robustness_runs.npz")
TS = [20, 30, 40, 50, 80, 100, 200, 300, 400, 500, 1000, 1500, 2000, 2500, 3000, 3500, 4000, 4500, 5000,
      5500, 6000, 6500, 7000, 7500]


def main():
    ys = np.stack([np.load(os.path.join(SRC, f"data-T8000-Run{j + 1}.npz"), allow_pickle=False)["data"][:max(TS)]
                   for j in range(100)]).astype(np.float32)
    tables = {}
    for method in ("extended", "cubature"):
        for form in ("sqrt", "std"):
            with np.load(os.path.join(SRC, f"ell_float32_{method}_{form}_100runs.npz"), allow_pickle=False) as d:
                tables[f"ell_{method}_{form}"] = np.asarray(d[d.files[0]], np.float32)
    np.savez_compressed(OUT, ys=ys, Ts=np.asarray(TS, np.int64), **tables)
    print(f"wrote {OUT}: ys {ys.shape}, tables {[(k, v.shape) for k, v in tables.items()]}")


if __name__ == "__main__":
    main()
