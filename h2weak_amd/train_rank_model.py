#!/usr/bin/env python3
"""Recipe of the shipped rank predictor rank_mlp_3d_rbf_tol10.npz.

    python -m h2weak_amd.train_rank_model harvest OUT.npz           (GPU: 16 assemblies)
    python -m h2weak_amd.train_rank_model harvest-test OUT.npz      (GPU: the two N = 262144 sets)
    python -m h2weak_amd.train_rank_model train IN.npz [TEST.npz]   (CPU or GPU, ~30 s)

Harvest: 3D RBF (choice 1), ACA tolerance 1e-10, n_max 64, points uniform in
[-1,1]^3 and N(0, 0.2^2) rejection-sampled into [-1,1]^3, N in {16384, 32768,
65536, 131072} with seeds 1 to 4 and N = 196608 with seeds 1 and 2 (about 200k
blocks).  Training: 200 epochs of Adam (lr 2e-3, cosine decay, batch 1024) on
mean squared log2(1 + rank) error, 10 % held out; parameters rounded to BF16.
The N = 262144 sets of the rank measurement (seed 7) are not part of the harvest.
"""
import sys

import numpy as np

from h2weak_amd import rank_mlp

SPECS = ([(k, N, s) for k in ("uniform", "gauss") for N in (16384, 32768, 65536, 131072) for s in (1, 2, 3, 4)]
         + [(k, 196608, s) for k in ("uniform", "gauss") for s in (1, 2)])
TEST_SPECS = [("uniform", 262144, 7), ("gauss", 262144, 7)]


def main():
    if len(sys.argv) < 3 or sys.argv[1] not in ("harvest", "harvest-test", "train"):
        print(__doc__)
        sys.exit(2)
    if sys.argv[1].startswith("harvest"):
        specs = SPECS if sys.argv[1] == "harvest" else TEST_SPECS
        F, R, S = rank_mlp.harvest(specs, log=lambda *a: print(*a, flush=True))
        np.savez(sys.argv[2], F=F, R=R, S=S)
        return
    with np.load(sys.argv[2], allow_pickle=False) as z:
        F, R = z["F"], z["R"]
    params, met = rank_mlp.train(F, R, epochs=200)
    meta = {"dim": 3, "choice": 1, "tol_pow": 10, "n_max": 64, "features": 16, "hidden": [64, 64, 64],
            "activation": "gelu (erf)", "params": "bf16", "target": "log2(1 + rank)",
            "train_sets": "uniform and gauss(0.2) in [-1,1]^3, N in {16384, 32768, 65536, 131072} seeds 1-4, "
                          "N = 196608 seeds 1 and 2",
            "train_samples": int(len(R)), **met}
    if len(sys.argv) > 3:
        # held-out sets: accuracy and the share of blocks whose rank exceeds the 1.2 x + 4 estimate
        with np.load(sys.argv[3], allow_pickle=False) as z:
            Ft, Rt, St = z["F"], z["R"], z["S"]
        pt = rank_mlp.forward(params, Ft)
        est = np.ceil(1.2 * np.maximum(0, np.exp2(pt) - 1)) + 4
        meta["test"] = {f"{k}_{N}_seed{sd}": {"mae_log2": float(np.mean(np.abs(pt[St == q] - np.log2(1 + Rt[St == q])))),
                                                "under_frac": float(np.mean(Rt[St == q] > est[St == q]))}
                        for q, (k, N, sd) in enumerate(TEST_SPECS)}
    rank_mlp.save(rank_mlp.default_model_path(3, 1, 10) or "rank_mlp_3d_rbf_tol10.npz", params, meta)
    print(meta)


if __name__ == "__main__":
    main()
