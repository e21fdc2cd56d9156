"""Host-side preparation done by the reference before its BFM operators run.

* kept_modes    validateModalInfo_ (src/bsa/BsaLibImpl.F90:545-599): modes whose
                max |phi| is not exactly 1 are dropped.
* build_psi     setPhitimesCLocalInstance_ (src/bsa/BsaLibImpl.F90:611-671):
                Psi(im, in, id) = sum_lib wfc(lib, id, in) * phi(dof(n, lib), modes(im)).
* skewness      computeSkewness_ (src/bsa/BsaLibImpl.F90:1630-1707).
* bkg_peak_width  validateAll_ (BsaLibImpl.F90:513-527) + computeBKGPeakWidths
                (src/structure/StructureImpl.F90:225-260): max over (component,
                direction) of U_ref/L, with L/U_ref maximised over wind zones first.
* damping_ratios  xsi_m = C_mm / (2 sqrt(K_m M_m)) of the kept modes.

These are O(NDOFS*NM) host loops (vectorised with numpy); everything O(NF^2)
runs in libbsahip.
"""
from __future__ import annotations

import numpy as np


def kept_modes(prob) -> np.ndarray:
    mx = np.abs(np.asarray(prob.phi)).max(axis=0)
    modes = (np.flatnonzero(mx == 1.0) + 1).astype(np.int32)
    if modes.size == 0:
        # same condition the reference aborts on (BsaLibImpl.F90:638-639)
        raise ValueError("Effective number of modes to be kept is 0 (no mode has max|phi| == 1)")
    return modes


def build_psi(prob, modes) -> np.ndarray:
    modes = np.asarray(modes, dtype=np.int64)
    nnl = int(np.size(prob.n_load))
    wfc = np.asarray(prob.wfc, dtype=np.float64)            # (NLIBSL, NDEGW, NNL)
    if wfc.shape[2] != nnl:
        raise ValueError("N. of loaded nodes does not match.")
    if wfc.shape[0] != np.size(prob.libs_load):
        raise ValueError("N. of loaded libs  does not match.")
    rows = (np.asarray(prob.n_load, dtype=np.int64)[:, None] - 1) * prob.nlibs + \
        (np.asarray(prob.libs_load, dtype=np.int64)[None, :] - 1)          # (NNL, NLIBSL)
    ph = np.asarray(prob.phi)[rows][:, :, modes - 1]                     # (NNL, NLIBSL, NM)
    # Psi(m, i, d) = sum_l wfc(l, d, i) * ph(i, l, m)
    psi = np.einsum("ldi,ilm->mid", wfc, ph)
    return np.asfortranarray(psi)


def skewness(m2, m3, nm: int) -> np.ndarray:
    """sk(m,n,p) = m3(m,n,p) / (sigma_m sigma_n sigma_p), sigma_m = sqrt(m2(m,m));
    flat index m + (n-1) nm + (p-1) nm^2 (as the reference)."""
    m2 = np.asarray(m2).reshape(nm, nm, order="F")
    sig = np.sqrt(np.diagonal(m2))
    den = sig[:, None, None] * sig[None, :, None] * sig[None, None, :]
    return (np.asarray(m3).reshape(nm, nm, nm, order="F") / den).reshape(-1, order="F")


def bkg_peak_width(prob, restrict: bool = False) -> float:
    L = np.asarray(prob.turb_scales, dtype=np.float64)            # (3, 3, NWZ)
    U = np.asarray(prob.u_mean_ref, dtype=np.float64)             # (NWZ,)
    v = (L / U[None, None, :]).max(axis=2)                        # (3, 3)
    w = np.where(v != 0.0, 1.0 / np.where(v != 0.0, v, 1.0), 0.0)
    # bsa_setMaxBkgPeakRestriction: only the first column (BsaMesherImpl.F90:298-302)
    return float(w[:, 0].max() if restrict else w.max())


def damping_ratios(prob, modes=None) -> np.ndarray:
    modes = kept_modes(prob) if modes is None else np.asarray(modes)
    mo = modes - 1
    c = np.diag(np.asarray(prob.Cm))[mo]
    return c / (2.0 * np.sqrt(np.asarray(prob.Km)[mo] * np.asarray(prob.Mm)[mo]))
