"""Write the golden vectors of tests/golden/*.npz from the REFERENCE itself.

Runs oracle/_ref/ref_harness (the reference's own Fortran BFM/BRM operators,
compiled from /root/reference by oracle/ref/build_ref.sh) on seeded synthetic
problems and stores inputs checksums and outputs as small fixtures:

* psd_f / psd_r      modal force / response PSD (getFM_full_tnm_vect_cls_ /
                     getRM_full_vect_cls_), full arrays
* bisp_f / bisp_r    modal force / response bispectra, full arrays for the
                     tiny cases, a fixed sample of grid points for config 1
* m2mf, m2mr, m3mf, m3mr   the quadrature of intgSpectraVect_ applied to the
                     reference's full spectra (the quadrature routine itself
                     lives in a reference unit that does not compile here and
                     is restated by the oracle; its weights are pinned by
                     tests/test_oracle.py)
* bfm_msh            getFM_full_tnm_scalar_msh_ at listed points (Mesher BFM,
                     exact S(f1+f2), including points whose f1+f2 is off-grid)
* *_diag             the ONLY_DIAG model (getFM_diag_tnlm_vect_cls_ /
                     getRM_diag_vect_cls_): psd (nf, nm), bispectra (full for
                     the tiny cases, the same sample for config 1) and the
                     quadrature m2mf/m2mr/m3mf/m3mr_diag (nm)

The reference run uses MALLOC_PERTURB_=255 (zero-filled allocations) so that
the never-initialised padding it reads for |f1+f2| > fmax holds the intended
zeros; a second run with MALLOC_PERTURB_=170 checks that exactly those points
depend on the padding (their count is stored as t2_points).

Usage (this container only; the GPU box never runs it):
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from bsalib_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent

from tests.cases import CASES  # noqa: E402


def _quad(prob, psd, bisp):
    f = prob.freqs()
    nm = int(round(psd.shape[1] ** 0.5))
    m2 = np.zeros(nm * nm)
    m3 = np.zeros(nm ** 3)
    O._intg(f, np.asfortranarray(psd), m2, np.asfortranarray(bisp), m3)
    return m2, m3


def _quad_diag(prob, psd, bisp):
    nm = psd.shape[1]
    m2, m3 = np.zeros(nm), np.zeros(nm)
    O._intg(prob.freqs(), np.asfortranarray(psd), m2, np.asfortranarray(bisp), m3)
    return m2, m3


def make(name: str):
    prob = CASES[name]()
    f = prob.freqs()
    nf = f.size
    rng = np.random.default_rng(123)
    # Mesher points: on-grid, off-grid, and f1+f2 beyond the Classic range
    fmax = f[-1]
    pts = rng.uniform(-fmax, fmax, size=(12, 2))
    pts = np.vstack([pts, [[fmax, fmax], [-fmax, 0.9 * fmax], [0.0, 0.0], [f[1], f[-2]]]])
    # zero-filled allocations: the padding of S(f1+f2) holds its intended zeros
    R = O.reference_run(prob, mode=7, fi=pts[:, 0], fj=pts[:, 1], malloc_perturb=255)
    # second run with 0xAA-filled allocations: points whose value moves read the padding
    R2 = O.reference_run(prob, mode=1, malloc_perturb=170)
    t2 = np.abs(R2["bisp_f"] - R["bisp_f"]).max(axis=2) > 0
    c = nf // 2
    ii, jj = np.meshgrid(np.arange(nf), np.arange(nf), indexing="ij")
    assert np.array_equal(t2, np.abs(ii + jj - 2 * c) > c), "padding reads outside |f1+f2| > fmax"
    m2f, m3f = _quad(prob, R["psd_f"], R["bisp_f"])
    m2r, m3r = _quad(prob, R["psd_r"], R["bisp_r"])
    out = dict(phi=prob.phi, natf=prob.natf, nod_corr_sum=np.array(prob.nod_corr.sum()),
               suvw=R["suvw"], psd_f=R["psd_f"], psd_r=R["psd_r"], m2mf=m2f, m2mr=m2r,
               m3mf=m3f, m3mr=m3r, msh_fi=pts[:, 0], msh_fj=pts[:, 1], bfm_msh=R["bfm_msh"],
               t2_points=np.array(int(t2.sum())))
    for k in ("f", "r"):
        m2d, m3d = _quad_diag(prob, R[f"psd_{k}_diag"], R[f"bisp_{k}_diag"])
        out[f"psd_{k}_diag"] = R[f"psd_{k}_diag"]
        out[f"m2m{k}_diag"], out[f"m3m{k}_diag"] = m2d, m3d
    if nf <= 20:
        out["bisp_f"] = R["bisp_f"]
        out["bisp_r"] = R["bisp_r"]
        out["bisp_f_diag"] = R["bisp_f_diag"]
        out["bisp_r_diag"] = R["bisp_r_diag"]
    else:
        i1 = rng.integers(0, nf, 300)
        i2 = rng.integers(0, nf, 300)
        i1[:4] = [0, nf - 1, 0, nf - 1]
        i2[:4] = [0, 0, nf - 1, nf - 1]
        out["sample_i1"], out["sample_i2"] = i1, i2
        out["bisp_f_sample"] = R["bisp_f"][i1, i2, :]
        out["bisp_r_sample"] = R["bisp_r"][i1, i2, :]
        out["bisp_f_diag_sample"] = R["bisp_f_diag"][i1, i2, :]
        out["bisp_r_diag_sample"] = R["bisp_r_diag"][i1, i2, :]
    np.savez_compressed(OUT / f"{name}.npz", **out)
    print(name, {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    if not O.reference_available():
        O.build()
        raise SystemExit("oracle/_ref/ref_harness missing: run oracle/ref/build_ref.sh first")
    for n in (sys.argv[1:] or CASES):
        make(n)
