"""Trains the ACA rank predictor on the GPU and writes its weights.

Training records are the ACA runs of H-matrix builds on three workloads,
swept over epsilon in {1e-4, 1e-6, 1e-8} and eta in {2, 5, 10}:
  * cylinder: N = 4000 on a cylinder of radius 1 and height 1, kernel
    1/(1e-5 + r), leaf size 10;
  * sphere: N = 1e5 on the unit sphere, 1/(4 pi r) with 1/(4 pi h) on the
    diagonal, leaf size 32;
  * spheroid: N = 2e5 on a 4:1 prolate spheroid (semi-axes 1, 1, 4), the same
    Laplace kernel, leaf size 32.
Usage: python -m htool_amd.train_predictor [--out htool_amd/data/rank_predictor.npz]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "rank_predictor.npz")


def workloads(scale: float = 1.0):
    from . import geometry as geo
    n_sph, n_spd = int(1e5 * scale), int(2e5 * scale)
    yield "cylinder", geo.cylinder(4000, seed=11), geo.KERNEL_INV_DELTA_R, 1e-5, 0.0, 10
    yield "sphere", geo.sphere(n_sph, seed=12), geo.KERNEL_LAPLACE, 0.0, \
        geo.laplace_diag(geo.sphere_spacing(n_sph)), 32
    pts = geo.ellipsoid(n_spd, (1.0, 1.0, 4.0), seed=13)
    h = np.sqrt(geo.prolate_spheroid_area(1.0, 4.0) / n_spd)
    yield "spheroid", pts, geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(h), 32


def collect(scale: float = 1.0, eps_list=(1e-4, 1e-6, 1e-8), eta_list=(2.0, 5.0, 10.0), log=print):
    from .hmatrix import Generator, HMatrix
    feats, terms, names = [], [], []
    for name, pts, kernel, delta, diag, mcs in workloads(scale):
        for eps in eps_list:
            for eta in eta_list:
                H = HMatrix(epsilon=eps, eta=eta, minclustersize=mcs)
                H.set_compression("partialACA")
                H.build(Generator(kernel, delta, diag), pts)
                f, t, _ = H.aca_records()
                feats.append(f)
                terms.append(t)
                names += [name] * len(t)
                log(f"  {name} eps={eps:g} eta={eta:g}: {len(t)} ACA records")
                del H
    return np.concatenate(feats), np.concatenate(terms), np.array(names)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--scale", type=float, default=1.0, help="scales the sphere and spheroid sizes")
    args = ap.parse_args(argv)
    from .predictor import RankPredictor, misprediction, targets_from_terms

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.time()
    feat, terms, names = collect(args.scale, log=log)
    t_collect = time.time() - t0
    rng = np.random.default_rng(args.seed)
    hold = rng.uniform(size=len(terms)) < 0.1
    P = RankPredictor(seed=args.seed)
    t0 = time.time()
    hist = P.train(feat[~hold], targets_from_terms(terms[~hold]), epochs=args.epochs, batch=args.batch, lr=args.lr,
                   seed=args.seed)
    t_train = time.time() - t0
    log(f"trained on {int((~hold).sum())} records in {t_train:.1f} s; loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    report = {"records": int(len(terms)), "train_records": int((~hold).sum()), "collect_s": round(t_collect, 2),
              "train_s": round(t_train, 2), "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
              "loss_first_epoch": float(hist[0]), "loss_last_epoch": float(hist[-1]), "held_out": {}}
    for name in np.unique(names):
        sel = hold & (names == name)
        m = misprediction(P.predict_terms(feat[sel]), terms[sel])
        report["held_out"][str(name)] = {"records": int(sel.sum()), **{k: round(v, 4) for k, v in m.items()}}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    P.save(args.out)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report))
    return 0


if __name__ == "__main__":
    sys.exit(main())
