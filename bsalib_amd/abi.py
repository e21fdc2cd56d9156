"""ctypes binding of libbsahip.so (C ABI declared in include/bsahip.h).

The library is loaded from bsalib_amd/_lib/ (built in-tree by
``bsalib_amd.build``).  A missing library or a missing GPU raises; nothing
here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import host

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libbsahip.so"

PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
QUAD_VECT, QUAD_RIEMANN, QUAD_TRAPZ = 0 << 4, 1 << 4, 2 << 4
SUM_TRUNCATED, SUM_EXACT = 0, 1 << 8
NO_PSD, NO_BISP = 1 << 9, 1 << 10
ONLY_DIAG = 1 << 11
FRF_PSD, FRF_GRID, FRF_POINTS = 0, 1, 2
EPI_GENERIC = 1 << 12          # testing: the generic quadrature epilogue only
DETERMINISTIC = 1 << 13        # fixed-order (bitwise reproducible) moment sums
# return codes (include/bsahip.h)
EINVAL, EDEVICE, ENOMEM, EUNSUPPORTED, ESTATE = -1, -2, -3, -4, -5

QUAD = {"vect": QUAD_VECT, "riemann": QUAD_RIEMANN, "trapezoid": QUAD_TRAPZ}
PREC = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}


class BsaHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bsahip error {code}: {msg}")
        self.code = code


class Problem_t(C.Structure):
    """Mirror of ``bsahip_problem`` (include/bsahip.h)."""
    _fields_ = [
        ("nn", C.c_int32), ("nnl", C.c_int32), ("nm", C.c_int32), ("ntc", C.c_int32),
        ("ndegw", C.c_int32), ("nwz", C.c_int32), ("psd_id", C.c_int32),
        ("pulsation", C.c_int32), ("ncorr_rows", C.c_int32), ("corr_col_mode", C.c_int32),
        ("pod_nmodes", C.c_int32), ("reserved", C.c_int32), ("pod_trunc_lim", C.c_double),
        ("n_load", C.POINTER(C.c_int32)), ("tc", C.POINTER(C.c_int32)),
        ("psi", C.POINTER(C.c_double)), ("nod_corr", C.POINTER(C.c_double)),
        ("u_node", C.POINTER(C.c_double)), ("wz_node", C.POINTER(C.c_int32)),
        ("turb_scales", C.POINTER(C.c_double)), ("turb_std", C.POINTER(C.c_double)),
        ("u_mean_ref", C.POINTER(C.c_double)), ("Mm", C.POINTER(C.c_double)),
        ("Km", C.POINTER(C.c_double)), ("Cm", C.POINTER(C.c_double)),
    ]


_lib = None

EXPORTS = ["bsahip_init", "bsahip_finalise", "bsahip_last_error", "bsahip_version",
           "bsahip_classic_run", "bsahip_classic_run_noalloc", "bsahip_classic_spectra",
           "bsahip_classic_bispectra_sel",
           "bsahip_bfm_points",
           "bsahip_last_kernel_time", "bsahip_plan_create", "bsahip_plan_run",
           "bsahip_plan_destroy", "bsahip_mesher_create", "bsahip_mesher_layout", "bsahip_mesher_ntiles",
           "bsahip_mesher_tiles", "bsahip_mesher_features", "bsahip_mesher_run",
           "bsahip_mesher_run_tiles", "bsahip_mesher_export", "bsahip_mesher_save",
           "bsahip_mesher_load", "bsahip_batch_create", "bsahip_batch_autotune",
           "bsahip_batch_run", "bsahip_batch_info", "bsahip_batch_destroy",
           "bsahip_mesher_destroy", "bsahip_modal_psd_points", "bsahip_modal_recombination",
           "bsahip_mesher_adapt", "bsahip_dgemm_strided_batched", "bsahip_mesher_run_batch",
           "bsahip_frf_apply", "bsahip_bfm_points_cls", "bsahip_mesher_adapt_run",
           "bsahip_diag_psd_points"]


def lib():
    """Load libbsahip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is missing: run `python -m bsalib_amd.build` "
                                    "(the HIP library is required, there is no CPU path)")
        L = C.CDLL(str(LIB_PATH))
        L.bsahip_last_error.restype = C.c_char_p
        L.bsahip_plan_create.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                         C.POINTER(C.c_void_p)]
        L.bsahip_plan_run.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 5
        L.bsahip_plan_destroy.argtypes = [C.c_void_p]
        L.bsahip_mesher_create.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]
        L.bsahip_mesher_ntiles.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.bsahip_mesher_layout.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                           C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.bsahip_mesher_tiles.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.bsahip_mesher_features.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
        L.bsahip_mesher_run.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.bsahip_mesher_run_tiles.argtypes = [C.c_void_p, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.bsahip_mesher_export.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int32,
                                           C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int32),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.bsahip_mesher_run_batch.argtypes = [C.c_int32, C.POINTER(C.c_void_p),
                                              C.POINTER(C.POINTER(C.c_int32)),
                                              C.POINTER(C.POINTER(C.c_double)),
                                              C.POINTER(C.POINTER(C.c_double)),
                                              C.POINTER(C.c_int64)]
        L.bsahip_mesher_save.argtypes = [C.c_void_p, C.c_char_p]
        L.bsahip_mesher_load.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_char_p,
                                         C.POINTER(C.c_void_p)]
        L.bsahip_mesher_destroy.argtypes = [C.c_void_p]
        L.bsahip_mesher_adapt.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int64)]
        L.bsahip_mesher_adapt_run.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                              C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise BsaHipError(rc, lib().bsahip_last_error().decode(errors="replace"))


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class Bound:
    """A problem packed into the C struct, keeping the numpy buffers alive."""

    def __init__(self, prob, modes=None, corr_col_mode: int = 0, psd_id=None,
                 pod_nmodes: int = 0, pod_trunc_lim: float = 0.0):
        """pod_nmodes / pod_trunc_lim: truncated POD of S(f) in the point BFM
        (bsahip_bfm_points and the Mesher); either needs corr_col_mode=1."""
        self.modes = host.kept_modes(prob) if modes is None else np.asarray(modes, np.int32)
        psi = host.build_psi(prob, self.modes)
        mo = self.modes - 1
        self.keep = dict(
            n_load=np.ascontiguousarray(prob.n_load, dtype=np.int32),
            tc=np.ascontiguousarray(prob.tc, dtype=np.int32),
            psi=np.asfortranarray(psi),
            nod_corr=np.asfortranarray(prob.nod_corr, dtype=np.float64),
            u_node=np.ascontiguousarray(prob.u_node, dtype=np.float64),
            wz_node=np.ascontiguousarray(prob.wz_node, dtype=np.int32),
            turb_scales=np.asfortranarray(prob.turb_scales, dtype=np.float64),
            turb_std=np.asfortranarray(prob.turb_std, dtype=np.float64),
            u_mean_ref=np.ascontiguousarray(prob.u_mean_ref, dtype=np.float64),
            Mm=np.ascontiguousarray(prob.Mm[mo], dtype=np.float64),
            Km=np.ascontiguousarray(prob.Km[mo], dtype=np.float64),
            Cm=np.ascontiguousarray(np.diag(prob.Cm)[mo], dtype=np.float64),
        )
        k = self.keep
        s = Problem_t()
        s.nn, s.nnl, s.nm, s.ntc = prob.nn, prob.nnl, len(self.modes), prob.ntc
        s.ndegw, s.nwz = psi.shape[2], prob.nwz
        s.psd_id = prob.psd_id if psd_id is None else psd_id
        s.pulsation = 1
        s.ncorr_rows = prob.nod_corr.shape[0]
        s.corr_col_mode = corr_col_mode
        s.pod_nmodes = pod_nmodes
        s.pod_trunc_lim = pod_trunc_lim
        for name in ("n_load", "tc", "wz_node"):
            setattr(s, name, _ptr(k[name], C.c_int32))
        for name in ("psi", "nod_corr", "u_node", "turb_scales", "turb_std", "u_mean_ref",
                     "Mm", "Km", "Cm"):
            setattr(s, name, _ptr(k[name], C.c_double))
        self.struct = s
        self.nm = len(self.modes)


def init(device: int = 0):
    _check(lib().bsahip_init(C.c_int(device)))


def classic_run(prob, quad: str = "vect", prec: str = "fp64", bound: Bound = None,
                psd: bool = True, bisp: bool = True, exact: bool = False,
                only_diag: bool = False, epi_generic: bool = False,
                deterministic: bool = False) -> dict:
    """Classic full-grid moments (bsahip_classic_run); exact=True evaluates
    S(f1+f2) beyond the grid instead of taking it as zero; only_diag=True runs
    the ONLY_DIAG model (every output then holds NM values); epi_generic=True
    sends every block of the plane quadrature through the generic epilogue."""
    b = bound or Bound(prob)
    nm = b.nm
    n2, n3 = (nm, nm) if only_diag else (nm * nm, nm ** 3)
    out = {k: np.zeros(n2) for k in ("m2mf", "m2mr", "m2o2mr")}
    out.update({k: np.zeros(n3) for k in ("m3mf", "m3mr")})
    flags = PREC[prec] | QUAD[quad] | (0 if psd else NO_PSD) | (0 if bisp else NO_BISP) | \
        (SUM_EXACT if exact else SUM_TRUNCATED) | (ONLY_DIAG if only_diag else 0) | \
        (EPI_GENERIC if epi_generic else 0) | (DETERMINISTIC if deterministic else 0)
    D = C.POINTER(C.c_double)
    _check(lib().bsahip_classic_run(C.byref(b.struct), C.c_int32(prob.nfreqs), C.c_double(prob.df),
                                    C.c_int32(flags), _ptr(out["m2mf"], C.c_double),
                                    _ptr(out["m2mr"], C.c_double), _ptr(out["m2o2mr"], C.c_double),
                                    _ptr(out["m3mf"], C.c_double), _ptr(out["m3mr"], C.c_double)))
    out["modes"] = b.modes
    return out


def classic_spectra(prob, exact: bool = False, bound: Bound = None) -> dict:
    """Modal PSDs (NF, NM^2) and bispectra (NF, NF, NM^3) on the Classic grid."""
    b = bound or Bound(prob)
    nm, nf = b.nm, prob.nf
    psd_f = np.zeros((nf, nm * nm), order="F")
    psd_r = np.zeros((nf, nm * nm), order="F")
    bfm = np.zeros((nf, nf, nm ** 3), order="F")
    brm = np.zeros((nf, nf, nm ** 3), order="F")
    flags = SUM_EXACT if exact else SUM_TRUNCATED
    _check(lib().bsahip_classic_spectra(C.byref(b.struct), C.c_int32(prob.nfreqs),
                                        C.c_double(prob.df), C.c_int32(flags),
                                        _ptr(psd_f, C.c_double), _ptr(psd_r, C.c_double),
                                        _ptr(bfm, C.c_double), _ptr(brm, C.c_double)))
    return {"psd_f": psd_f, "psd_r": psd_r, "bisp_f": bfm, "bisp_r": brm}


def classic_bispectra_sel(prob, sel, exact: bool = False, only_diag: bool = False,
                          bound: Bound = None, prec: str = "fp64") -> dict:
    """Bispectra of the selected modal entries on the Classic grid, (NF, NF,
    len(sel)); sel is 1-based over the kept modes, m + (n-1) NM + (p-1) NM^2
    (the mode position with only_diag).  prec "fp32" / "bf16x3": the G values
    with that G kernel's operand rounding.  bsahip_classic_bispectra_sel."""
    b = bound or Bound(prob)
    nf = prob.nf
    s = np.ascontiguousarray(np.atleast_1d(sel), dtype=np.int32)
    bfm = np.zeros((nf, nf, s.size), order="F")
    brm = np.zeros((nf, nf, s.size), order="F")
    flags = (SUM_EXACT if exact else SUM_TRUNCATED) | PREC[prec]
    if only_diag:
        flags |= ONLY_DIAG
    _check(lib().bsahip_classic_bispectra_sel(C.byref(b.struct), C.c_int32(prob.nfreqs),
                                              C.c_double(prob.df), C.c_int32(flags),
                                              C.c_int32(s.size), _ptr(s, C.c_int32),
                                              _ptr(bfm, C.c_double), _ptr(brm, C.c_double)))
    return {"bisp_f": bfm, "bisp_r": brm}


def bfm_points(prob, fi, fj, with_brm: bool = True, bound: Bound = None, only_diag: bool = False):
    """Modal bispectra at arbitrary points, S(f1+f2) exact (bsahip_bfm_points):
    (NM^3, npts), or (NM, npts) for the diagonal model."""
    b = bound or Bound(prob)
    fi = np.ascontiguousarray(fi, dtype=np.float64)
    fj = np.ascontiguousarray(fj, dtype=np.float64)
    if fj.size != fi.size:
        raise ValueError("one fj per fi")
    n = fi.size
    ne = b.nm if only_diag else b.nm ** 3
    bfm = np.zeros((ne, n), order="F")
    brm = np.zeros((ne, n), order="F") if with_brm else None
    _check(lib().bsahip_bfm_points(C.byref(b.struct), C.c_int32(n), _ptr(fi, C.c_double),
                                   _ptr(fj, C.c_double), C.c_int32(ONLY_DIAG if only_diag else 0),
                                   _ptr(bfm, C.c_double), _ptr(brm, C.c_double)))
    return bfm, brm


def bfm_points_cls(prob, fi, fj, with_brm: bool = True, bound: Bound = None,
                   only_diag: bool = False):
    """Modal bispectra at points of the Classic grid of prob (nfreqs, df) with
    its truncation, S(|f|) = 0 beyond (nfreqs - 1) df (bsahip_bfm_points_cls):
    (NM^3, npts), or (NM, npts) for the diagonal model."""
    b = bound or Bound(prob)
    fi = np.ascontiguousarray(fi, dtype=np.float64)
    fj = np.ascontiguousarray(fj, dtype=np.float64)
    if fj.size != fi.size:
        raise ValueError("one fj per fi")
    n = fi.size
    ne = b.nm if only_diag else b.nm ** 3
    bfm = np.zeros((ne, n), order="F")
    brm = np.zeros((ne, n), order="F") if with_brm else None
    _check(lib().bsahip_bfm_points_cls(C.byref(b.struct), C.c_int32(prob.nfreqs),
                                       C.c_double(prob.df), C.c_int32(ONLY_DIAG if only_diag else 0),
                                       C.c_int32(n), _ptr(fi, C.c_double), _ptr(fj, C.c_double),
                                       _ptr(bfm, C.c_double), _ptr(brm, C.c_double)))
    return bfm, brm


def diag_psd_points(prob, f, bound: Bound = None):
    """Modal force PSD of the diagonal model at frequencies f (bsahip_diag_psd_points):
    array (NM, npts)."""
    b = bound or Bound(prob)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    out = np.zeros((b.nm, f.size), order="F")
    _check(lib().bsahip_diag_psd_points(C.byref(b.struct), C.c_int32(f.size), _ptr(f, C.c_double),
                                        _ptr(out, C.c_double)))
    return out


def frf_apply(prob, kind: int, data, f1, f2=None, flags: int = 0, bound: Bound = None):
    """FRF weighting of modal spectra (bsahip_frf_apply), returned as a new
    array: kind FRF_PSD data (n1, NM^2); FRF_GRID data (n1, n2, NM^3);
    FRF_POINTS data (NM^3, n1) at the points (f1[q], f2[q]); NM in place of
    NM^2 and NM^3 with flags=ONLY_DIAG."""
    b = bound or Bound(prob)
    f1 = np.ascontiguousarray(f1, dtype=np.float64).reshape(-1)
    f2 = None if f2 is None else np.ascontiguousarray(f2, dtype=np.float64).reshape(-1)
    out = np.array(data, dtype=np.float64, order="F", copy=True)
    n2 = 0 if f2 is None else f2.size
    nm = b.nm
    e2, e3 = (nm, nm) if flags & ONLY_DIAG else (nm * nm, nm ** 3)
    want = {FRF_PSD: f1.size * e2, FRF_GRID: f1.size * n2 * e3, FRF_POINTS: f1.size * e3}.get(kind)
    if want is not None and out.size != want:
        raise ValueError(f"data holds {out.size} values, kind {kind} needs {want}")
    if kind == FRF_POINTS and n2 != f1.size:
        raise ValueError("FRF_POINTS needs one f2 per f1")
    _check(lib().bsahip_frf_apply(C.byref(b.struct), C.c_int32(flags), C.c_int32(kind),
                                  C.c_int64(f1.size), _ptr(f1, C.c_double), C.c_int64(n2),
                                  _ptr(f2, C.c_double), _ptr(out, C.c_double)))
    return out


def modal_psd_points(prob, f, bound: Bound = None):
    """Modal force PSD S_Q(n, p; f) at arbitrary frequencies
    (bsahip_modal_psd_points): array (nm, nm, npts)."""
    b = bound or Bound(prob)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    out = np.zeros((b.nm * b.nm, f.size), order="F")
    _check(lib().bsahip_modal_psd_points(C.byref(b.struct), C.c_int32(f.size), _ptr(f, C.c_double),
                                         _ptr(out, C.c_double)))
    return out.reshape(b.nm, b.nm, f.size, order="F")


def modal_recombination(phi, m2, m3, m2o2=None, ccc: bool = True):
    """Nodal moments from modal ones (bsahip_modal_recombination): phi (ndofs, nm),
    m2/m2o2 (nm*nm), m3 (nm^3).  Returns dict m2, m3, m2o2 (or None), sk."""
    phi = np.asfortranarray(phi, dtype=np.float64)
    ndofs, nm = phi.shape
    m2 = np.ascontiguousarray(m2, dtype=np.float64).reshape(-1)
    m3 = np.ascontiguousarray(m3, dtype=np.float64).reshape(-1)
    m22 = None if m2o2 is None else np.ascontiguousarray(m2o2, dtype=np.float64).reshape(-1)
    if m2.size != nm * nm or m3.size != nm ** 3 or (m22 is not None and m22.size != nm * nm):
        raise ValueError("m2/m2o2 need nm*nm and m3 nm^3 entries")
    o = {k: np.zeros(ndofs) for k in ("m2", "m3", "sk")}
    o["m2o2"] = np.zeros(ndofs) if m22 is not None else None
    _check(lib().bsahip_modal_recombination(
        C.c_int32(ndofs), C.c_int32(nm), _ptr(phi, C.c_double), _ptr(m2, C.c_double),
        _ptr(m3, C.c_double), _ptr(m22, C.c_double), C.c_int32(1 if ccc else 0),
        _ptr(o["m2"], C.c_double), _ptr(o["m3"], C.c_double), _ptr(o["m2o2"], C.c_double),
        _ptr(o["sk"], C.c_double)))
    return o


def last_kernel_ms():
    ms = C.c_double(0)
    n = C.c_int32(0)
    _check(lib().bsahip_last_kernel_time(C.byref(ms), C.byref(n)))
    return ms.value, n.value


class Plan:
    """A Classic run resident on the device (bsahip_plan_*): create once, run
    many times.  The problem's host arrays are only read by the constructor."""

    def __init__(self, prob, quad: str = "vect", prec: str = "fp64", bound: Bound = None,
                 psd: bool = True, bisp: bool = True, deterministic: bool = False):
        b = bound or Bound(prob)
        self.nm = b.nm
        self.modes = b.modes
        flags = PREC[prec] | QUAD[quad] | (0 if psd else NO_PSD) | (0 if bisp else NO_BISP) | \
            (DETERMINISTIC if deterministic else 0)
        h = C.c_void_p()
        _check(lib().bsahip_plan_create(C.addressof(b.struct), C.c_int32(prob.nfreqs),
                                        C.c_double(prob.df), C.c_int32(flags), C.byref(h)))
        self.h = h

    def run(self) -> dict:
        nm = self.nm
        out = {k: np.zeros(nm * nm) for k in ("m2mf", "m2mr", "m2o2mr")}
        out.update({k: np.zeros(nm ** 3) for k in ("m3mf", "m3mr")})
        _check(lib().bsahip_plan_run(self.h, *(_ptr(out[k], C.c_double) for k in
                                               ("m2mf", "m2mr", "m2o2mr", "m3mf", "m3mr"))))
        out["modes"] = self.modes
        return out

    def close(self):
        if self.h:
            lib().bsahip_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class MesherCfg(C.Structure):
    """Mirror of ``bsahip_mesher_cfg`` (include/bsahip.h)."""
    _fields_ = [("bkg_base_rfmnt", C.c_int32), ("tile_cells", C.c_int32),
                ("diag_crest", C.c_int32), ("brm_cells", C.c_int32), ("bkg_peakw", C.c_double),
                ("bkg_area_ext", C.c_double), ("peak_area_ext", C.c_double),
                ("max_area_ext", C.c_double),
                ("natf", C.POINTER(C.c_double)), ("xsi", C.POINTER(C.c_double))]


def mesher_cfg(prob, bkg_base_rfmnt: int = 20, tile_cells: int = 32, bkg_peakw=None,
               bkg_area_ext: float = 2.0, peak_area_ext: float = 2.0, max_area_ext: float = 2.0,
               modes=None, diag_crest: bool = False, natf=None, xsi=None, brm_cells: int = 0):
    """bsahip_mesher_cfg for a problem; returns (cfg, arrays kept alive).  natf
    and xsi (kept modes) default to the problem's; they only shape the layout.
    brm_cells > 0: uniform BRM spacing df_ref / brm_cells on every tile."""
    modes = host.kept_modes(prob) if modes is None else np.asarray(modes, np.int32)
    mo = modes - 1
    natf = np.ascontiguousarray(np.asarray(prob.natf, dtype=np.float64)[mo] if natf is None
                                else np.asarray(natf, dtype=np.float64))
    xsi = np.ascontiguousarray(host.damping_ratios(prob, modes) if xsi is None
                               else np.asarray(xsi, dtype=np.float64))
    cfg = MesherCfg()
    cfg.bkg_base_rfmnt, cfg.tile_cells = bkg_base_rfmnt, tile_cells
    cfg.diag_crest, cfg.brm_cells = int(diag_crest), int(brm_cells)
    cfg.bkg_peakw = host.bkg_peak_width(prob) if bkg_peakw is None else float(bkg_peakw)
    cfg.bkg_area_ext, cfg.peak_area_ext, cfg.max_area_ext = bkg_area_ext, peak_area_ext, max_area_ext
    cfg.natf = _ptr(natf, C.c_double)
    cfg.xsi = _ptr(xsi, C.c_double)
    return cfg, (natf, xsi)


def mesher_layout(cfg, nm: int):
    """Tile layout only (host computation): geom (ntiles, 10), rule levels."""
    n = C.c_int32(0)
    _check(lib().bsahip_mesher_layout(C.addressof(cfg), C.c_int32(nm), C.c_int32(0), None, None,
                                      C.byref(n)))
    geom = np.zeros((n.value, 10))
    rule = np.zeros(n.value, dtype=np.int32)
    _check(lib().bsahip_mesher_layout(C.addressof(cfg), C.c_int32(nm), n, _ptr(geom, C.c_double),
                                      _ptr(rule, C.c_int32), C.byref(n)))
    return geom, rule


class Mesher:
    """Zone Mesher resident on the device (bsahip_mesher_*).  The constructor
    lays out the tiles and evaluates every tile's level-0 BFM (pre-mesh);
    run() integrates with the given per-tile refinement levels (post-mesh)."""

    def __init__(self, prob, bkg_base_rfmnt: int = 20, tile_cells: int = 32, bkg_peakw=None,
                 bkg_area_ext: float = 2.0, peak_area_ext: float = 2.0, max_area_ext: float = 2.0,
                 bound: Bound = None, diag_crest: bool = False, natf=None, xsi=None,
                 load: str = None, only_diag: bool = False, brm_cells: int = 0):
        """load: a checkpoint written by save() for the same problem and
        settings, read instead of evaluating the pre-mesh (bsahip_mesher_load).
        only_diag: the ONLY_DIAG model (NM entries per point instead of NM^3)."""
        b = bound or Bound(prob)
        self.nm = b.nm
        self.nc = b.nm if only_diag else b.nm ** 3
        self.modes = b.modes
        flags = C.c_int32(ONLY_DIAG if only_diag else 0)
        cfg, self._keep = mesher_cfg(prob, bkg_base_rfmnt, tile_cells, bkg_peakw, bkg_area_ext,
                                     peak_area_ext, max_area_ext, modes=b.modes,
                                     diag_crest=diag_crest, natf=natf, xsi=xsi,
                                     brm_cells=brm_cells)
        self.cfg = cfg
        self.bkg_peakw = cfg.bkg_peakw
        h = C.c_void_p()
        self.h = None
        if load is None:
            _check(lib().bsahip_mesher_create(C.addressof(b.struct), C.addressof(cfg), flags,
                                              C.byref(h)))
        else:
            _check(lib().bsahip_mesher_load(C.addressof(b.struct), C.addressof(cfg), flags,
                                            str(load).encode(), C.byref(h)))
        self.h = h
        n = C.c_int32(0)
        _check(lib().bsahip_mesher_ntiles(self.h, C.byref(n)))
        self.ntiles = n.value

    def tiles(self):
        geom = np.zeros((self.ntiles, 10))
        rule = np.zeros(self.ntiles, dtype=np.int32)
        _check(lib().bsahip_mesher_tiles(self.h, _ptr(geom, C.c_double), _ptr(rule, C.c_int32)))
        return geom, rule

    def features(self, ns: int = 16) -> np.ndarray:
        out = np.zeros((self.ntiles, ns * ns))
        _check(lib().bsahip_mesher_features(self.h, C.c_int32(ns), _ptr(out, C.c_double)))
        return out

    def run(self, levels=None, per_tile: bool = False) -> dict:
        """Post-mesh with per-tile levels (None: rule-based); per_tile adds
        each tile's m3mr contribution as "tile_m3mr" (ntiles, entries)."""
        nm3 = self.nc
        m3mf, m3mr = np.zeros(nm3), np.zeros(nm3)
        counts = np.zeros(3, dtype=np.int64)
        lv = None
        if levels is not None:
            lv = self._levels(levels)
        tiles = np.zeros((self.ntiles, nm3)) if per_tile else None
        _check(lib().bsahip_mesher_run_tiles(self.h, _ptr(lv, C.c_int32), _ptr(tiles, C.c_double),
                                             _ptr(m3mf, C.c_double), _ptr(m3mr, C.c_double),
                                             _ptr(counts, C.c_int64)))
        out = {"m3mf": m3mf, "m3mr": m3mr, "evals_pre": int(counts[0]),
               "evals_post": int(counts[1]), "brm_points": int(counts[2]), "modes": self.modes}
        if per_tile:
            out["tile_m3mr"] = tiles
        return out

    def adapt(self, tol: float = 1e-3):
        """A-posteriori refinement levels for a target m3mr error tol relative to
        max|m3mr| (bsahip_mesher_adapt): (levels, BFM evaluations of those levels)."""
        lv = np.zeros(self.ntiles, dtype=np.int32)
        ev = C.c_int64(0)
        _check(lib().bsahip_mesher_adapt(self.h, C.c_double(tol), _ptr(lv, C.c_int32), C.byref(ev)))
        return lv, int(ev.value)

    def adapt_run(self, tol: float = 1e-3) -> dict:
        """adapt() and the run at its levels in one call (bsahip_mesher_adapt_run):
        the dict of run() plus "levels" and "evals"."""
        lv = np.zeros(self.ntiles, dtype=np.int32)
        ev = C.c_int64(0)
        m3mf, m3mr = np.zeros(self.nc), np.zeros(self.nc)
        counts = np.zeros(3, dtype=np.int64)
        _check(lib().bsahip_mesher_adapt_run(self.h, C.c_double(tol), _ptr(lv, C.c_int32),
                                             C.byref(ev), _ptr(m3mf, C.c_double),
                                             _ptr(m3mr, C.c_double), _ptr(counts, C.c_int64)))
        return {"m3mf": m3mf, "m3mr": m3mr, "evals_pre": int(counts[0]),
                "evals_post": int(counts[1]), "brm_points": int(counts[2]), "modes": self.modes,
                "levels": lv, "evals": int(ev.value)}

    def _levels(self, levels):
        if levels is None:
            return None
        return np.ascontiguousarray(np.broadcast_to(np.asarray(levels, dtype=np.int32),
                                                    (self.ntiles,)))

    def export(self, levels=None, coef=None) -> dict:
        """BRM at every tile's post-mesh points (bsahip_mesher_export): coef
        (entries, ncol) combinations, None for every entry.  Returns dims
        (ntiles, 2), fi, fj (npts) and val (npts, ncol)."""
        lv = self._levels(levels)
        nm3 = self.nc
        dims = np.zeros((self.ntiles, 2), dtype=np.int32)
        n = C.c_int64(0)
        L = lib()
        _check(L.bsahip_mesher_export(self.h, _ptr(lv, C.c_int32), 0, None, 0,
                                      _ptr(dims, C.c_int32), None, None, None, C.byref(n)))
        if coef is not None:
            coef = np.asfortranarray(np.asarray(coef, dtype=np.float64).reshape(nm3, -1, order="F"))
        ncol = nm3 if coef is None else coef.shape[1]
        fi, fj = np.zeros(n.value), np.zeros(n.value)
        val = np.zeros((n.value, ncol))
        _check(L.bsahip_mesher_export(self.h, _ptr(lv, C.c_int32), ncol, _ptr(coef, C.c_double),
                                      n.value, _ptr(dims, C.c_int32), _ptr(fi, C.c_double),
                                      _ptr(fj, C.c_double), _ptr(val, C.c_double), C.byref(n)))
        return {"dims": dims, "fi": fi, "fj": fj, "val": val}

    def save(self, path):
        """Checkpoint of the resident zones (bsahip_mesher_save)."""
        _check(lib().bsahip_mesher_save(self.h, str(path).encode()))

    def close(self):
        if self.h:
            lib().bsahip_mesher_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass



def mesher_run_batch(meshers, levels=None) -> list:
    """bsahip_mesher_run_batch: the post-mesh of several Meshers at once, with
    per-Mesher levels (each None for the rule-based ones).  One dict per Mesher
    with m3mf, m3mr and the counts of Mesher.run."""
    n = len(meshers)
    hs = (C.c_void_p * n)(*[m.h for m in meshers])
    lv = [None if levels is None or levels[i] is None else meshers[i]._levels(levels[i])
          for i in range(n)]
    LP = C.POINTER(C.c_int32)
    larr = (LP * n)(*[_ptr(v, C.c_int32) if v is not None else LP() for v in lv])
    outs = [(np.zeros(m.nc), np.zeros(m.nc)) for m in meshers]
    DP = C.POINTER(C.c_double)
    f = (DP * n)(*[_ptr(o[0], C.c_double) for o in outs])
    r = (DP * n)(*[_ptr(o[1], C.c_double) for o in outs])
    counts = np.zeros(3 * n, dtype=np.int64)
    _check(lib().bsahip_mesher_run_batch(C.c_int32(n), hs, larr, f, r, _ptr(counts, C.c_int64)))
    return [{"m3mf": outs[i][0], "m3mr": outs[i][1], "evals_pre": int(counts[3 * i]),
             "evals_post": int(counts[3 * i + 1]), "brm_points": int(counts[3 * i + 2]),
             "modes": meshers[i].modes} for i in range(n)]

class Batch:
    """Structures of uneven size resident on the device (bsahip_batch_*):
    run(mode=0) is one persistent G launch over all of them replayed from a
    HIP graph, run(mode=1) one G launch per structure."""

    def __init__(self, probs, quad: str = "vect", prec: str = "fp64", psd: bool = True):
        L = lib()
        if not getattr(L, "_batch_bound", False):
            L.bsahip_batch_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_int32),
                                              C.POINTER(C.c_double), C.c_int32, C.c_int32,
                                              C.POINTER(C.c_void_p)]
            L.bsahip_batch_autotune.argtypes = [C.c_void_p, C.POINTER(C.c_int32),
                                                C.POINTER(C.c_int32)]
            L.bsahip_batch_run.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
            L.bsahip_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
            L.bsahip_batch_destroy.argtypes = [C.c_void_p]
            L._batch_bound = True
        self.bounds = [Bound(p) for p in probs]
        self.nm = [b.nm for b in self.bounds]
        self.evals = int(sum((2 * p.nfreqs - 1) ** 2 for p in probs))
        n = len(probs)
        ptrs = (C.c_void_p * n)(*[C.addressof(b.struct) for b in self.bounds])
        nfh = np.array([p.nfreqs for p in probs], dtype=np.int32)
        df = np.array([p.df for p in probs], dtype=np.float64)
        flags = PREC[prec] | QUAD[quad] | (0 if psd else NO_PSD)
        h = C.c_void_p()
        self.h = None
        _check(L.bsahip_batch_create(ptrs, _ptr(nfh, C.c_int32), _ptr(df, C.c_double), n, flags,
                                     C.byref(h)))
        self.h = h

    def autotune(self):
        m, c = C.c_int32(0), C.c_int32(0)
        _check(lib().bsahip_batch_autotune(self.h, C.byref(m), C.byref(c)))
        return m.value, c.value

    def run(self, mode: int = 0) -> dict:
        n = len(self.nm)
        out = {k: [np.zeros(nm ** 3) for nm in self.nm] for k in ("m3mf", "m3mr")}
        out["m2mr"] = [np.zeros(nm ** 2) for nm in self.nm]
        arr = {k: (C.c_void_p * n)(*[v.ctypes.data for v in out[k]]) for k in out}
        _check(lib().bsahip_batch_run(self.h, C.c_int32(mode), arr["m3mf"], arr["m3mr"], arr["m2mr"]))
        return out

    def info(self) -> dict:
        ms, items = C.c_double(0), C.c_int64(0)
        m, c = C.c_int32(0), C.c_int32(0)
        _check(lib().bsahip_batch_info(self.h, C.byref(ms), C.byref(items), C.byref(m), C.byref(c)))
        return {"ms_step": ms.value, "items": items.value, "grid_mult": m.value, "chunk": c.value}

    def close(self):
        if self.h:
            lib().bsahip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass
