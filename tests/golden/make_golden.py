"""Writes the committed golden fixtures of the oracle tests.

* fft_numpy.npz -- seeded complex inputs and their numpy.fft.fft (pocketfft,
  FP64) outputs for lengths 8 .. 4096; they pin the oracle's FP64 FFT.
* reference_max_error.csv -- the max pointwise relative error (percent, the
  reference's "%lf" column) that the reference recorded for its FMM-FFT on the
  ramp input x_n = n+1 with B=32 points per leaf and T=15 terms, one row per
  (N, P), read from the reference's
  plotting/raw_data/cluster_output_error.csv (data only).

Run from the repository root: python tests/golden/make_golden.py
"""
import csv
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_CSV = "/root/reference/plotting/raw_data/cluster_output_error.csv"


def fft_vectors():
    rng = np.random.default_rng(20240601)
    out = {}
    for n in (8, 16, 64, 256, 1024, 4096):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x{n}"] = x
        out[f"X{n}"] = np.fft.fft(x)
    np.savez(os.path.join(HERE, "fft_numpy.npz"), **out)


def reference_errors():
    rows = {}
    with open(REF_CSV) as f:
        for line in f:
            parts = [p.strip() for p in line.split(",")]
            if len(parts) != 6:
                continue
            try:
                n, p, b, t = (int(v) for v in parts[:4])
                float(parts[4])
                err = parts[5]
                float(err)
            except ValueError:
                continue
            rows.setdefault((n, p), (b, t, err))
    with open(os.path.join(HERE, "reference_max_error.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["N", "P", "B", "T", "max_rel_err_pct"])
        for (n, p) in sorted(rows):
            b, t, err = rows[(n, p)]
            w.writerow([n, p, b, t, err])


if __name__ == "__main__":
    fft_vectors()
    if os.path.exists(REF_CSV):
        reference_errors()
