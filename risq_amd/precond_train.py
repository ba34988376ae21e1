"""Training of the learned Davidson preconditioner (run on the GPU).

Training systems follow the recipe of the synthetic molecular integrals
(h_pp = -1 + 2p/15, h_pq ~ N(0, 0.05^2), (pq|rs) from K = 2 n_orb Cholesky
vectors with L ~ N(0, 0.08^2)) over energy-truncated determinant spaces: the
small systems at n_orb 20-28 and 8-10 electrons (seeds 100 + i), and smaller
spaces (1.5e5 determinants) of the n_orb 40, 12-electron system at seeds 2-5
(the 2e6-determinant benchmark system is seed 1 and is not trained on).  For
each system the library converges Davidson tightly, then re-runs it with the
restart that keeps the previous Ritz vector (the fast path's) and records, at
every iteration, per-determinant features with the diagonal correction and
the exact correction (hci_mlp_collect).  The network is trained on the pooled rows with
the library's BF16-MFMA kernels and saved as an .npz (parameters + feature
standardisation).

    python -m risq_amd.precond_train [out.npz] [--recipe small]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

from .dets import energy_truncated_space
from .fcidump import synthetic_molecular
from .hci import HciContext

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "precond_mlp.npz")
# trained on the small systems only (n_orb 20-28, 4-5 electrons per spin)
SMALL_RECIPE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "precond_mlp_small.npz")

# (n_orb, electrons per spin, determinants, seed)
TRAINING_SYSTEMS = [(20, 4, 30000, 100), (22, 5, 30000, 101), (24, 4, 40000, 102),
                    (26, 5, 40000, 103), (28, 4, 50000, 104), (28, 5, 50000, 105),
                    (40, 6, 150000, 2), (40, 6, 150000, 3), (40, 6, 150000, 4), (40, 6, 150000, 5)]


def collect(ctx: HciContext, systems=TRAINING_SYSTEMS, per_iter: int = 4096, log=print):
    Xs, ms, ys = [], [], []
    for n_orb, ne, n_det, seed in systems:
        t = time.time()
        ints = synthetic_molecular(n_orb, 2 * n_orb, seed)
        eps = np.array([ints.h1(p, p) for p in range(n_orb)])
        dets = energy_truncated_space(eps, ne, ne, n_det)
        ctx.set_integrals(ints)
        h = ctx.build_h(dets, ctx.ham_diag(dets))
        exact = ctx.davidson(h, tol=1e-8, etol=1e-12, max_iters=2000)
        X, m, y = ctx.mlp_collect(h, exact.coeffs, per_iter=per_iter, cap=per_iter * 400, max_iters=400,
                                  tol=1e-6, etol=1e-10, restart_keep=2)
        Xs.append(X)
        ms.append(m)
        ys.append(y)
        log(f"  system n_orb={n_orb} {ne}a{ne}b n={dets.shape[0]} nnz_u={h.nnz} E0={exact.energy:.8f} "
            f"rows={X.shape[0]} ({time.time() - t:.1f}s)")
        h.free()
    return np.concatenate(Xs), np.concatenate(ms), np.concatenate(ys)


def train(ctx: HciContext, X, m, y, epochs: int = 12, batch: int = 4096, lr: float = 1e-3, seed: int = 7):
    ctx.mlp_init(seed)
    return ctx.mlp_train(X, m, y, epochs=epochs, batch=batch, lr=lr, seed=seed)


def save(ctx: HciContext, path: str, meta: dict) -> None:
    p, mu, sd = ctx.mlp_get_params()
    np.savez(path, params=p, feat_mean=mu, feat_std=sd, **{k: np.array(v) for k, v in meta.items()})


def load(ctx: HciContext, path: str = DEFAULT_PATH) -> None:
    with np.load(path, allow_pickle=False) as z:
        ctx.mlp_set_params(z["params"], z["feat_mean"], z["feat_std"])


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("out", nargs="?", default=None)
    ap.add_argument("--recipe", choices=["all", "small"], default="all",
                    help="small: only the n_orb 20-28, 8-10 electron systems (the recipe of the learned-"
                         "preconditioner workload); all: those plus four n_orb-40 spaces")
    args = ap.parse_args(sys.argv[1:] if argv is None else argv)
    path = args.out or (SMALL_RECIPE_PATH if args.recipe == "small" else DEFAULT_PATH)
    systems = [s for s in TRAINING_SYSTEMS if s[0] <= 28] if args.recipe == "small" else TRAINING_SYSTEMS
    ctx = HciContext(0)
    t = time.time()
    X, m, y = collect(ctx, systems)
    print(f"collected {X.shape[0]} rows in {time.time() - t:.1f}s", flush=True)
    st = train(ctx, X, m, y)
    print(f"trained: loss {st['first_epoch_loss']:.4f} -> {st['last_epoch_loss']:.4f} in {st['seconds']:.1f}s",
          flush=True)
    base = float(np.mean((m - y) ** 2))
    save(ctx, path, {"rows": X.shape[0], "loss_first": st["first_epoch_loss"], "loss_last": st["last_epoch_loss"],
                     "loss_diagonal": base})
    print(f"diagonal-correction loss on the same rows: {base:.4f}; saved {path}", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
