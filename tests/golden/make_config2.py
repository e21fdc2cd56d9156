"""Write tests/golden/config2.npz: the full-size tower of configuration 2
(NN=200 loaded nodes, u/v/w turbulence, NM=10 modes, Classic grid of 513 x 513
points over [-4, 4] Hz) through the CPU oracle.

The reference's own vectorised Classic BFM cannot run at this size (its triple
node loop makes NTC*NNL^3*NF^2*NM^3 = 6e15 element updates), so the fixture
comes from the oracle's factorised form, which tests/test_oracle.py pins to the
reference harness's full arrays and to the literal triple loop:

* m2mf, m2mr, m2o2mr   modal PSD moments (getFM_full_tnm_vect_cls_ PSD part,
                       getRM_full_vect_cls_, intgSpectraVect_)
* m3mf, m3mr           third moments under intgSpectraVect_ (corner weight -1/2)
                       with S(f1+f2) zero beyond the grid (the Classic path)
* sample_i1/i2         grid indexes of 300 sampled points: the four corners,
                       points with |f1+f2| > fmax (truncated S(f1+f2)), the
                       axes and random interior points
* sel                  27 selected 1-based entries m + (n-1)NM + (p-1)NM^2
                       (modes {1, 5, 10} in every position)
* bisp_f_sample / bisp_r_sample (300, 27)  force / response bispectra there

Usage (this container; the GPU box only reads the fixture):
    python tests/golden/make_config2.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from bsalib_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent / "config2.npz"


def sample_points(nf: int, rng) -> tuple[np.ndarray, np.ndarray]:
    c = (nf - 1) // 2
    pts = [(0, 0), (0, nf - 1), (nf - 1, 0), (nf - 1, nf - 1), (c, c), (c, 0), (0, c),
           (nf - 1, c), (c, nf - 1)]
    # beyond the f1+f2 edge (|F1 + F2| > c) on both sides
    while len(pts) < 80:
        i, j = rng.integers(0, nf, size=2)
        if abs(i + j - 2 * c) > c:
            pts.append((int(i), int(j)))
    # on the edge |F1 + F2| = c
    for i in rng.integers(0, nf, size=20):
        j = 3 * c - i if i >= c else c - i
        if 0 <= j < nf:
            pts.append((int(i), int(j)))
    while len(pts) < 300:
        i, j = rng.integers(0, nf, size=2)
        pts.append((int(i), int(j)))
    a = np.array(pts[:300], dtype=np.int64)
    return a[:, 0], a[:, 1]


def main() -> None:
    p = synth.config2()
    nf = p.nf
    assert nf == 513, nf
    rng = np.random.default_rng(2024)
    i1, i2 = sample_points(nf, rng)
    full = O.classic_factorised(p, rule="vect", truncate=True, nthreads=0)
    smp = O.classic_factorised(p, rule="vect", truncate=True, nthreads=0, want_bisp=True,
                               point_sample=(i1, i2))
    m = O.modal_psd(p)
    nm = len(full["modes"])
    modes = np.array([0, 4, 9])
    sel = (modes[:, None, None] + nm * modes[None, :, None] + nm * nm * modes[None, None, :])
    sel = sel.ravel(order="F").astype(np.int32)            # 0-based
    np.savez_compressed(
        OUT, m2mf=m["m2mf"], m2mr=m["m2mr"], m2o2mr=m["m2o2mr"], m3mf=full["m3mf"],
        m3mr=full["m3mr"], sample_i1=i1, sample_i2=i2, sel=sel + 1,
        bisp_f_sample=np.ascontiguousarray(smp["bfm"][sel, :].T),
        bisp_r_sample=np.ascontiguousarray(smp["brm"][sel, :].T))
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
