"""Learned refinement proposer of the zone Mesher.

The reference refines every Mesher zone by a fixed rule (delta validation,
src/bsa/meshing/zones/_shared_poly2d.fi:735-795): the BFM sampling must resolve
the narrowest resonance peak of the zone's interest modes, whether or not the
zone's contribution to the third moment needs it.  Here a small MLP (4 linear
layers, width 128, GELU; bsahip_mlp_* kernels) looks at each tile after the
level-0 pass and proposes level 8 (refine fully) or level 1 (keep the level-0
samples), 2 or 4 in between.  It is trained on tiles of seeded structures
whose per-tile contributions at every level the Mesher itself computes; the
target is the cheapest set of levels (greedy in error per evaluation) that
keeps the third-moment error bound within budget, encoded ordinally as
log2(level)/3 in one logit.

Features per tile (tile_features):
  * tile centre and size relative to the domain half-width, log2 of its
    level-0 cells per axis (6);
  * the interval policy of each axis, one-hot (10);
  * for each resonance family -- H_m(f1): the f1 range, H_n(f2): the f2 range,
    H_p(f1+f2): the f1+f2 range -- the three modes nearest to the range (in
    units of their peak width w_n = 2 xsi_n f_n peak_area_ext): log10(1 + d/w),
    f_n / half-width, 100 xsi_n (27);
  * ||B(f1,f2)||_2 over mode triples on a 16 x 16 stencil of the level-0
    samples, log10 relative to the structure's largest value (256), plus the
    tile's log-maximum and log-range (2).
"""
from __future__ import annotations

import math
import time
from pathlib import Path

import numpy as np

from . import abi, host, synth
from .mlp import MLP

NS = 16
POLICIES = (3, 4, 5, 9, 100)
NFEAT = 6 + 10 + 27 + NS * NS + 2
WEIGHTS = Path(__file__).resolve().parent / "data" / "mesher_proposer.npz"


def mesher_settings(prob) -> dict:
    """Mesher settings of the proposer workloads: base (background) zone of 33
    points per side (bkg_base_rfmnt), tiles of at most 8 x 8 level-0 cells,
    background width U_ref / L_u (the along-wind turbulence peak)."""
    L = np.asarray(prob.turb_scales)[0, 0, :]
    U = np.asarray(prob.u_mean_ref)
    return dict(bkg_base_rfmnt=33, tile_cells=8, bkg_peakw=float(np.max(U / L)))


def tile_features(prob, mesher: abi.Mesher, geom=None, stencil=None) -> np.ndarray:
    if geom is None:
        geom, _ = mesher.tiles()
    if stencil is None:
        stencil = mesher.features(NS)
    modes = mesher.modes
    mo = modes - 1
    fn = np.asarray(prob.natf, dtype=np.float64)[mo]
    xi = host.damping_ratios(prob, modes)
    wn = 2.0 * xi * fn * 2.0                          # peak_area_ext = 2
    F = float(np.max(np.abs(geom[:, :4])))
    nt = len(geom)
    X = np.zeros((nt, NFEAT), dtype=np.float64)
    a1, b1, a2, b2 = geom[:, 0], geom[:, 1], geom[:, 2], geom[:, 3]
    X[:, 0] = 0.5 * (a1 + b1) / F
    X[:, 1] = 0.5 * (a2 + b2) / F
    X[:, 2] = (b1 - a1) / F
    X[:, 3] = (b2 - a2) / F
    X[:, 4] = np.log2(geom[:, 4])
    X[:, 5] = np.log2(geom[:, 5])
    for k, pol in enumerate(POLICIES):
        X[:, 6 + k] = geom[:, 6] == pol
        X[:, 11 + k] = geom[:, 7] == pol
    col = 16
    for lo, hi in ((a1, b1), (a2, b2), (a1 + a2, b1 + b2)):
        # distance of [lo, hi] to +-f_n, in peak widths
        d = np.minimum(_dist(lo, hi, fn), _dist(lo, hi, -fn)) / wn[None, :]
        order = np.argsort(d, axis=1)[:, :3]
        for r in range(3):
            n = order[:, r] if d.shape[1] > r else order[:, -1]
            X[:, col] = np.log10(1.0 + d[np.arange(nt), n])
            X[:, col + 1] = fn[n] / F
            X[:, col + 2] = 100.0 * xi[n]
            col += 3
    s = np.asarray(stencil, dtype=np.float64)
    G = max(float(s.max()), 1e-300)
    ls = np.log10(s / G + 1e-12)
    X[:, col:col + NS * NS] = ls
    col += NS * NS
    X[:, col] = ls.max(axis=1)
    X[:, col + 1] = ls.max(axis=1) - ls.min(axis=1)
    assert col + 2 == NFEAT
    return X.astype(np.float32)


def _dist(lo, hi, f):
    lo = lo[:, None]
    hi = hi[:, None]
    f = f[None, :]
    return np.where(f < lo, lo - f, np.where(f > hi, f - hi, 0.0))


LEVELS = (1, 2, 4, 8)


def tile_errors(tl: np.ndarray, t8: np.ndarray) -> np.ndarray:
    """Per tile: max over mode triples of |level-8 - level-l contribution|,
    relative to max |converged m3mr|."""
    ref = np.abs(t8.sum(axis=0)).max()
    return np.abs(t8 - tl).max(axis=1) / ref


def tile_costs(geom) -> np.ndarray:
    """(ntiles, 4): new BFM evaluations of a tile at levels 1, 2, 4, 8."""
    n1, n2 = geom[:, 4:5], geom[:, 5:6]
    L = np.array(LEVELS, dtype=np.float64)[None, :]
    return (n1 * L + 1) * (n2 * L + 1) - (n1 + 1) * (n2 + 1)


def greedy_levels(err: np.ndarray, cost: np.ndarray, budget: float) -> np.ndarray:
    """Levels (index into LEVELS) per tile: starting from level 1 everywhere,
    upgrade the tile whose move buys the most error per evaluation until the
    summed per-tile errors (a bound on the third-moment error) fit the
    budget.  err, cost: (ntiles, 4) with err[:, 3] = 0."""
    import heapq
    nt = err.shape[0]
    lvl = np.zeros(nt, dtype=np.int64)
    total = float(err[:, 0].sum())
    heap = []

    def push(t):
        c = lvl[t]
        best = None
        for k in range(c + 1, 4):
            gain = err[t, c] - err[t, k]
            dc = cost[t, k] - cost[t, c]
            if gain > 0:
                r = gain / max(dc, 1.0)
                if best is None or r > best[0]:
                    best = (r, k)
        if best is not None:
            heapq.heappush(heap, (-best[0], t, best[1], c))

    for t in range(nt):
        push(t)
    while total > budget and heap:
        _, t, k, c = heapq.heappop(heap)
        if lvl[t] != c:
            continue
        total -= err[t, c] - err[t, k]
        lvl[t] = k
        push(t)
    return lvl


def training_structure(seed: int):
    """Seeded deck or tower for training: small decks, towers, and decks of
    the size class of the evaluation deck (NN 150-300, NM 14-20)."""
    rng = np.random.default_rng(10_000 + seed)
    u = rng.random()
    if u < 0.45:
        nm = int(rng.integers(6, 13))
        return synth.deck(nn=int(rng.integers(40, 121)), nm=nm, seed=seed,
                          fmin=float(rng.uniform(0.08, 0.3)), fmax=float(rng.uniform(1.0, 2.5)),
                          xi=float(rng.uniform(0.004, 0.02)), u10=float(rng.uniform(20, 45)))
    if u < 0.7:
        nm = int(rng.integers(3, 7))
        natf = np.sort(rng.uniform(0.15, 2.0, nm))
        xi = rng.uniform(0.005, 0.02, nm)
        return synth.tower(nn=int(rng.integers(20, 61)), nm=nm, natf=natf, xi=xi, seed=seed,
                           zone_per_node=False)
    return synth.deck(nn=int(rng.integers(150, 301)), nm=int(rng.integers(14, 21)), seed=seed,
                      fmin=float(rng.uniform(0.08, 0.15)), fmax=float(rng.uniform(1.5, 2.5)),
                      xi=float(rng.uniform(0.004, 0.01)), u10=float(rng.uniform(20, 45)))


def validation_structure(seed: int):
    """Held-out decks of the evaluation size (NN 400-500, NM 18-20)."""
    rng = np.random.default_rng(50_000 + seed)
    return synth.deck(nn=int(rng.integers(400, 501)), nm=int(rng.integers(18, 21)),
                      seed=1000 + seed, fmin=float(rng.uniform(0.08, 0.12)),
                      fmax=float(rng.uniform(1.8, 2.2)), xi=float(rng.uniform(0.004, 0.006)),
                      u10=float(rng.uniform(25, 35)))


def structure_samples(prob, budget: float = 5e-4):
    """Features, ordinal targets log2(level)/3 of the greedy levels, and the
    per-tile errors and costs of one structure."""
    M = abi.Mesher(prob, **mesher_settings(prob))
    try:
        X = tile_features(prob, M)
        geom, _ = M.tiles()
        runs = [M.run(levels=l, per_tile=True)["tile_m3mr"] for l in LEVELS]
        err = np.stack([tile_errors(r, runs[3]) for r in runs], axis=1)
        cost = tile_costs(geom)
        lvl = greedy_levels(err, cost, budget)
        y = (lvl / 3.0).astype(np.float32)
        info = dict(ntiles=M.ntiles, refine_frac=float((lvl > 0).mean()))
    finally:
        M.close()
    return X, y, err, cost, info


def build_dataset(nstruct: int, seed0: int = 0, log=print):
    Xs, ys, errs, costs, groups = [], [], [], [], []
    t0 = time.time()
    for s in range(seed0, seed0 + nstruct):
        p = training_structure(s)
        X, y, err, cost, info = structure_samples(p)
        Xs.append(X)
        ys.append(y)
        errs.append(err)
        costs.append(cost)
        groups.append(np.full(len(y), s, dtype=np.int32))
        log(f"[proposer] structure {s}: {p.name} NN={p.nn} NM={p.nm} tiles {info['ntiles']} "
            f"refine {info['refine_frac']:.3f} ({time.time() - t0:.0f} s)")
    return (np.concatenate(Xs), np.concatenate(ys), np.concatenate(errs), np.concatenate(costs),
            np.concatenate(groups))


def train(X, y, cost, epochs: int = 40, seed: int = 0, log=print):
    """Trains the 4 x 128 GELU network on the device with soft binary
    cross-entropy on the ordinal target y = log2(level)/3.  Tiles that need
    refinement are weighted up by the ratio of unrefined to refined tiles."""
    mean = X.mean(axis=0)
    sd = X.std(axis=0)
    inv = np.where(sd > 1e-6, 1.0 / np.maximum(sd, 1e-6), 0.0)
    pos = max(float((y > 0).mean()), 1e-3)
    w = np.where(y > 0, 0.5 / pos, 0.5 / (1 - pos)).astype(np.float32)
    net = MLP(X.shape[1], 128, 4, seed=seed)
    net.set(mean=mean, inv_scale=inv)
    losses = net.train(X, y, w, epochs=epochs, batch=512, lr=1e-3, weight_decay=1e-5, seed=seed)
    log(f"[proposer] training loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    return net, mean, inv, losses


def levels_from_logits(z: np.ndarray, shift: float) -> np.ndarray:
    """Ordinal decoding: p = sigmoid(z - shift); level 1, 2, 4, 8 for p below
    1/6, 1/2, 5/6 and above (the midpoints between the targets 0, 1/3, 2/3, 1)."""
    p = 1.0 / (1.0 + np.exp(-(np.asarray(z, dtype=np.float64) - shift)))
    k = (p >= 1 / 6).astype(int) + (p >= 0.5) + (p >= 5 / 6)
    return np.asarray(LEVELS, dtype=np.int32)[k]


def save(net: MLP, mean, inv, threshold: float, path: Path = WEIGHTS, meta=None):
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=net.params(), mean=np.asarray(mean, np.float32),
             inv=np.asarray(inv, np.float32), nin=net.nin, width=net.width, depth=net.depth,
             threshold=np.float32(threshold), meta=np.asarray(str(meta or {})))


class Proposer:
    """Trained network + decision threshold (logit) loaded from an .npz."""

    def __init__(self, path: Path = WEIGHTS):
        with np.load(path, allow_pickle=False) as z:
            self.nin, self.width, self.depth = int(z["nin"]), int(z["width"]), int(z["depth"])
            self.threshold = float(z["threshold"])
            self.net = MLP(self.nin, self.width, self.depth)
            self.net.set(z["params"], z["mean"], z["inv"])

    def logits(self, prob, mesher: abi.Mesher) -> np.ndarray:
        return self.net.forward(tile_features(prob, mesher))

    def levels(self, prob, mesher: abi.Mesher, threshold=None) -> np.ndarray:
        """Refinement level per tile; threshold (a logit shift, default the
        stored one) trades evaluations against error."""
        th = self.threshold if threshold is None else threshold
        return levels_from_logits(self.logits(prob, mesher), th)


def sweep(prob, net: MLP, thresholds, log=print, target: float = 1e-3, plain: bool = True) -> dict:
    """Converged (every tile at level 8) and MLP-guided Mesher runs of one
    structure for several decision thresholds, and (plain=True) every plain
    setting: uniform levels 1, 2, 4, the rule-based levels and the a-posteriori
    adaptive levels for `target` (bsahip_mesher_adapt).  BFM evaluations and
    third-moment errors against the converged m3mr."""
    M = abi.Mesher(prob, **mesher_settings(prob))
    try:
        t0 = time.perf_counter()
        z = net.forward(tile_features(prob, M))
        t_inf = time.perf_counter() - t0
        conv = M.run(levels=8)
        ref = np.abs(conv["m3mr"]).max()

        def err(r):
            return float(np.abs(r["m3mr"] - conv["m3mr"]).max() / ref)

        def evals(r):
            return r["evals_pre"] + r["evals_post"]
        out = dict(name=f"{prob.name} NN={prob.nn} NM={prob.nm}", ntiles=M.ntiles,
                   evals_conv=evals(conv), features_and_inference_s=t_inf, mlp={}, plain={})
        if plain:
            for lv in (1, 2, 4):
                r = M.run(levels=lv)
                out["plain"][f"uniform_{lv}"] = dict(evals=evals(r), err=err(r))
            r = M.run()
            out["plain"]["rule"] = dict(evals=evals(r), err=err(r))
            t1 = time.perf_counter()
            r = M.adapt_run(target)
            out["plain"]["adaptive"] = dict(evals=r["evals"], err=err(r), tol=target,
                                            run_s=time.perf_counter() - t1)
            out["plain"]["uniform_8"] = dict(evals=evals(conv), err=0.0)
            ok = {k: v for k, v in out["plain"].items() if v["err"] <= target}
            best = min(ok, key=lambda k: ok[k]["evals"])
            out["cheapest_plain"] = dict(name=best, **ok[best])
            out["evals_rule"], out["err_rule"] = out["plain"]["rule"]["evals"], out["plain"]["rule"]["err"]
        for th in thresholds:
            t1 = time.perf_counter()
            r = M.run(levels=levels_from_logits(z, th))
            out["mlp"][th] = dict(evals=evals(r), err=err(r), run_s=time.perf_counter() - t1)
    finally:
        M.close()
    return out


def evaluate(prob, proposer: Proposer, log=print, target: float = 1e-3) -> dict:
    """The stored threshold only (see sweep): the MLP's evaluations and error,
    every plain setting, and the reduction against the cheapest plain setting
    that meets the target."""
    r = sweep(prob, proposer.net, [proposer.threshold], target=target)
    m = r.pop("mlp")[proposer.threshold]
    r.update(evals_mlp=m["evals"], err_mlp=m["err"])
    r["reduction_vs_conv"] = r["evals_conv"] / r["evals_mlp"]
    r["reduction_vs_rule"] = r["evals_rule"] / r["evals_mlp"]
    r["reduction_vs_cheapest_plain"] = r["cheapest_plain"]["evals"] / r["evals_mlp"]
    log(f"[proposer] {r}")
    return r


THRESHOLDS = (8.0, 6.0, 5.0, 4.0, 3.0, 2.0, 1.0, 0.0, -1.0, -2.0, -4.0)


def select_threshold(net: MLP, nval: int, target: float, thresholds=THRESHOLDS, log=print):
    """The most economical threshold whose errors on the first nval held-out
    decks all stay within half the target (the last one if none does)."""
    res = []
    for s in range(nval):
        r = sweep(validation_structure(s), net, thresholds, plain=False)
        log(f"[proposer] validation {r['name']}: converged {r['evals_conv']}; MLP " +
            ", ".join(f"{th:g}: {m['evals']} / {m['err']:.1e}" for th, m in r["mlp"].items()))
        res.append(r)
    for th in thresholds:
        if max(r["mlp"][th]["err"] for r in res) <= 0.5 * target:
            return th, res
    return thresholds[-1], res


def learning_curve(sizes, nval: int, epochs: int, target: float, log=print,
                   on_row=None, save_largest: Path = None) -> list:
    """One training set of max(sizes) structures; for each size n, a network
    trained on the tiles of the first n structures, its threshold chosen on
    nval held-out decks, and its evaluations and error on the configuration-5
    deck against the cheapest plain setting that meets the target.  on_row(rows)
    is called after each size; save_largest keeps the largest size's network."""
    nmax = max(sizes)
    X, y, err, cost, grp = build_dataset(nmax, 0, log=log)
    rows = []
    for n in sorted(sizes):
        sel = grp < n
        net, mean, inv, losses = train(X[sel], y[sel], cost[sel], epochs=epochs, log=log)
        th, _ = select_threshold(net, nval, target, log=log)
        r = sweep(synth.config5(), net, [th], target=target)
        m = r["mlp"][th]
        rows.append(dict(structures=n, tiles=int(sel.sum()), refined=float((y[sel] > 0).mean()),
                         loss=float(losses[-1]), threshold=th, evals_mlp=m["evals"],
                         err_mlp=m["err"], cheapest_plain=r["cheapest_plain"],
                         reduction_vs_cheapest_plain=r["cheapest_plain"]["evals"] / m["evals"]))
        log(f"[proposer] curve {rows[-1]}")
        if on_row:
            on_row(rows)
        if save_largest and n == nmax:
            save(net, mean, inv, th, Path(save_largest),
                 meta=dict(structures=n, epochs=epochs, samples=int(sel.sum()),
                           refined=rows[-1]["refined"], loss0=float(losses[0]),
                           loss=float(losses[-1]), validation=nval))
            log(f"[proposer] saved {save_largest} threshold {th}")
        net.close()
    return rows


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train the Mesher proposer on the GPU")
    ap.add_argument("--structures", type=int, default=200)
    ap.add_argument("--val", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--target", type=float, default=1e-3)
    ap.add_argument("--curve", default="",
                    help="comma-separated training-set sizes: write a learning curve (JSON, to "
                         "--out, after each size) instead of a network")
    ap.add_argument("--save-largest", default="",
                    help="with --curve: also save the largest size's network there")
    args = ap.parse_args()
    abi.init(0)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if args.curve:
        def write(rows):
            Path(args.out).write_text(json.dumps(dict(target=args.target, validation=args.val,
                                                      epochs=args.epochs, rows=rows), indent=1))
        learning_curve([int(v) for v in args.curve.split(",")], args.val, args.epochs,
                       args.target, log=log, on_row=write,
                       save_largest=Path(args.save_largest) if args.save_largest else None)
        log(f"[proposer] wrote {args.out}")
        return
    X, y, err, cost, grp = build_dataset(args.structures, 0, log=log)
    log(f"[proposer] {len(y)} tiles, {float((y > 0).mean()):.4f} refined")
    net, mean, inv, losses = train(X, y, cost, epochs=args.epochs, log=log)
    best, _ = select_threshold(net, args.val, args.target, log=log)
    save(net, mean, inv, best, Path(args.out),
         meta=dict(structures=args.structures, epochs=args.epochs, samples=int(len(y)),
                   refined=float((y > 0).mean()), loss0=float(losses[0]),
                   loss=float(losses[-1]), validation=args.val))
    log(f"[proposer] saved {args.out} threshold {best}")


if __name__ == "__main__":
    main()
