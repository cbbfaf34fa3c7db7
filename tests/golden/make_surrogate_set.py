#!/usr/bin/env python3
"""Writes tests/golden/surrogate_train.npz: the training set of the spin-up
surrogate, labelled by the FP64 device physics (needs the GPU and the built
library).

2000 river basins, the 100 largest of each of twenty glb_15min-shaped 1440x720
networks (land within +-70 degrees; seeds 500..503, five networks each), every
basin with its own runoff climatology (cmf.surrogate.basins_from_networks). Each
group of 500 is spun up as one batched FP64 handle from the physical prior to an
annual storage change below 1e-5 (cmf.surrogate.equilibrium_labels). Of the
cells the network corrects (memory < 60 days), 30% are drawn at random, then at
most 120 per basin, so every basin is represented and the file stays small.

Stored per cell: the 24 features (float16), the target log(equilibrium storage /
prior storage) (float32), the equilibrium and prior storages (float32) and the
basin index (int16).
Usage: python tests/golden/make_surrogate_set.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "cama-flood_v4_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

from cmf import surrogate as sg  # noqa: E402

PER_BASIN = 120


def main():
    Xs, Ys, Ss, Ps, Bs = [], [], [], [], []
    off = 0
    for chunk in range(4):
        basins = sg.basins_from_networks(5, 100, seed=500 + chunk, nx=1440, ny=720, max_lat=70.0)
        X, Y, res = sg.equilibrium_labels(basins, tol=1e-5, max_years=40, log=lambda s: print("  " + s, flush=True))
        ps = np.concatenate([sg.prior(b.m, b.clim).storage for b in basins])
        bid = np.concatenate([np.full(b.m.n, off + k, np.int32) for k, b in enumerate(basins)])
        off += len(basins)
        sto = res.state["rivsto"] + res.state["fldsto"]
        rng = np.random.default_rng(chunk)
        keep = sg.fast_cells(X) & (rng.random(X.shape[0]) < 0.3)
        Xs.append(X[keep].astype(np.float16))
        Ys.append(Y[keep, 0].astype(np.float32))
        Ss.append(sto[keep].astype(np.float32))
        Ps.append(ps[keep].astype(np.float32))
        Bs.append(bid[keep].astype(np.int16))
        print(f"group {chunk}: {len(basins)} basins, kept {int(keep.sum())} cells", flush=True)
    X, Y, S, P, B = (np.concatenate(a) for a in (Xs, Ys, Ss, Ps, Bs))
    rng = np.random.default_rng(1)
    order = np.argsort(B, kind="stable")
    nb = int(B.max()) + 1
    lo = np.searchsorted(B[order], np.arange(nb))
    hi = np.searchsorted(B[order], np.arange(nb), side="right")
    keep = np.zeros(B.size, bool)
    for b in range(nb):
        idx = order[lo[b]:hi[b]]
        if idx.size > PER_BASIN:
            idx = rng.choice(idx, PER_BASIN, replace=False)
        keep[idx] = True
    np.savez_compressed(os.path.join(HERE, "surrogate_train.npz"), X=X[keep], Y0=Y[keep], sto=S[keep],
                        prior=P[keep], bid=B[keep])
    print(f"surrogate_train.npz: {int(keep.sum())} cells of {nb} basins")


if __name__ == "__main__":
    main()
