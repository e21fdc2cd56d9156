"""CPU FP64 oracle for the BsaLib BFM hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module, and only as the checker / the timed CPU baseline.  The
product package ``bsalib_amd`` never imports it.

numpy restatements here (small, closed-form pieces) and ctypes bindings to
``bsa_oracle.c`` (the loop-heavy restatements).  Every function cites the
reference routine it follows.  How each function is pinned independently of
its own formula is listed in oracle/README.md.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libbsa_oracle.so"
TWO_PI = 2.0 * np.pi
FOUR_PI = 4.0 * np.pi

_lib = None


def build(force: bool = False) -> Path:
    """Compile bsa_oracle.c with gcc (OpenMP) into oracle/_build/."""
    src = HERE / "bsa_oracle.c"
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < src.stat().st_mtime:
        LIB_PATH.parent.mkdir(exist_ok=True)
        subprocess.check_call(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c11",
                               "-D_DEFAULT_SOURCE", str(src), "-o", str(LIB_PATH), "-lm"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(str(LIB_PATH))
    return _lib


def _p(a, ct=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64F(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


# ---------------------------------------------------------------------------
# wind PSDs -- src/wind/WindPSDImpl.F90
# ---------------------------------------------------------------------------
def eval_psd(prob, f, tc: int, idir: int = 1, pulsation: bool = True) -> np.ndarray:
    """evalPSD_ (WindPSDImpl.F90:123-144): PSD(nf, NNL) for component ``tc``
    (1-based), divided by 4*pi under the pulsation convention (:143)."""
    f = np.asarray(f, dtype=np.float64).reshape(-1)
    nl = prob.n_load - 1
    wz = prob.wz_node[nl] - 1
    U = prob.u_node[nl]
    sig = prob.turb_std[tc - 1, wz]
    pid = prob.psd_id
    if pid == 1:      # vonKarmanPSD_ (:149-226)
        if tc == 1:
            Lr = prob.turb_scales[0, idir - 1, wz]
            LU = Lr / U
            x = np.abs(f)[:, None] * LU[None, :]
            den = (1.0 + 70.7 * x * x) ** (5.0 / 6.0)
            psd = 4.0 * (LU * sig ** 2)[None, :] / den
        else:
            Lr = prob.turb_scales[tc - 1, idir - 1, wz]
            d = 2.0 * f[:, None] * (Lr / U)[None, :]
            d = d * d
            r2 = (1.0 + 70.7 * d) ** (11.0 / 6.0) * U[None, :]
            r3 = Lr[None, :] * (1.0 + 188.4 * d) / r2 * (sig ** 2)[None, :]
            psd = 4.0 * r3
    elif pid == 2:    # kaimalPSD_ returns zero (:231-243)
        psd = np.zeros((f.size, nl.size))
    elif pid == 3:    # davenportPSD_Greisch_ (:249-277)
        cst = 0.65 * 1200.0
        um = prob.u_mean_ref[wz]
        psd = (cst * sig ** 2 / um)[None, :] / \
            (1.0 + (cst * f[:, None] / um[None, :]) ** 2.0) ** (5.0 / 6.0)
    elif pid == 5:    # davenportPSD_Uliege_ (:283-316)
        Lr = prob.turb_scales[tc - 1, idir - 1, wz]
        cLU = Lr / U
        cF = np.abs(f)[:, None] * cLU[None, :]
        psd = (2.0 / 3.0) * cF * cLU[None, :] * (sig ** 2)[None, :] / (1.0 + cF ** 2) ** (4.0 / 3.0)
    else:
        raise ValueError(f"unsupported psd id {pid}")
    if pulsation:
        psd = psd / FOUR_PI
    return np.asfortranarray(psd)


def suvw_table(prob, f) -> np.ndarray:
    """S_uvw(NF, NNL*NTC) as mainClassic_ builds it (BsaClassicImpl.F90:103-119)."""
    return np.asfortranarray(np.concatenate([eval_psd(prob, f, int(t)) for t in prob.tc], axis=1))


# ---------------------------------------------------------------------------
# modal data -- src/bsa/BsaLibImpl.F90
# ---------------------------------------------------------------------------
def kept_modes(prob) -> np.ndarray:
    """validateModalInfo_ (BsaLibImpl.F90:545-599): keep modes with max|phi| == 1."""
    mx = np.abs(prob.phi).max(axis=0)
    return (np.nonzero(mx == 1.0)[0] + 1).astype(np.int32)


def psi(prob, modes=None) -> np.ndarray:
    """setPhitimesCLocalInstance_ (BsaLibImpl.F90:611-671):
    Psi(im, in, id) = sum_lib wfc(lib, id, in) * phi((n-1)*NLIBS + libs_load(lib), modes(im))."""
    if modes is None:
        modes = kept_modes(prob)
    nm = len(modes)
    nnl = prob.nnl
    ndegw = prob.wfc.shape[1]
    out = np.zeros((nm, nnl, ndegw), order="F")
    for a in range(nnl):
        n = int(prob.n_load[a])
        rows = (n - 1) * prob.nlibs + prob.libs_load - 1
        ph = prob.phi[np.ix_(rows, np.asarray(modes) - 1)]          # (NLIBSL, nm)
        for d in range(ndegw):
            out[:, a, d] = (prob.wfc[:, d, a][:, None] * ph).sum(axis=0)
    return out


def frf(prob, f, modes=None):
    """H_m(f) = 1/(K - w^2 M + i w C_mm), w = 2 pi f (functionsImpl.F90:1641-1651)."""
    if modes is None:
        modes = kept_modes(prob)
    mo = np.asarray(modes) - 1
    w = TWO_PI * np.asarray(f, dtype=np.float64)[..., None]
    r = -(w * w * prob.Mm[mo]) + prob.Km[mo]
    i = w * np.diag(prob.Cm)[mo]
    d = r * r + i * i
    return r / d - 1j * i / d


def skewness(m2: np.ndarray, m3: np.ndarray, nm: int) -> np.ndarray:
    """computeSkewness_ (BsaLibImpl.F90:1630-1707): sk(m,n,p) = m3/(s_m s_n s_p)
    with s_m = sqrt(m2(m,m)); index m + (n-1)nm + (p-1)nm^2."""
    sig = np.sqrt(np.asarray(m2).reshape(nm, nm, order="F").diagonal())
    den = sig[:, None, None] * sig[None, :, None] * sig[None, None, :]
    return np.asarray(m3).reshape(nm, nm, nm, order="F") / den


# ---------------------------------------------------------------------------
# Classic vectorised path -- literal restatement (tiny sizes only)
# ---------------------------------------------------------------------------
def classic_vect(prob, with_bisp: bool = True, nthreads: int = 1) -> dict:
    """mainClassic_ vectorised branch (BsaClassicImpl.F90:137-212):
    getFM_full_tnm_vect_cls_ -> intgSpectraVect_ -> getRM_full_vect_cls_ ->
    intgSpectraVect_ -> psd*w^2 -> intgSpectraVect_."""
    L = lib()
    f = prob.freqs()
    nf = f.size
    modes = kept_modes(prob)
    nm = len(modes)
    P = psi(prob, modes)
    S = suvw_table(prob, f)
    psd = np.zeros((nf, nm * nm), order="F")
    bisp = np.zeros((nf, nf, nm ** 3), order="F") if with_bisp else None
    L.orc_bfm_vect_cls(C.c_int(nf), _p(f), C.c_int(prob.nnl), C.c_int(prob.nn), C.c_int(nm),
                       C.c_int(prob.ntc), _p(_i32(prob.tc), C.c_int),
                       _p(_i32(prob.n_load), C.c_int), _p(P), C.c_int(P.shape[2]),
                       _p(_f64F(prob.nod_corr)), C.c_long(prob.nod_corr.shape[0]), _p(S),
                       _p(psd), _p(bisp), C.c_int(nthreads))
    out = {"f": f, "psd_f": psd.copy(), "bisp_f": None if bisp is None else bisp.copy()}
    m2 = np.zeros(nm * nm)
    m3 = np.zeros(nm ** 3)
    _intg(f, psd, m2, bisp, m3)
    out["m2mf"], out["m3mf"] = m2.copy(), m3.copy()
    L.orc_brm_vect_cls(C.c_int(nf), _p(f), C.c_int(nm), _p(modes, C.c_int),
                       C.c_int(prob.phi.shape[1]), _p(_f64F(prob.Mm)), _p(_f64F(prob.Km)),
                       _p(_f64F(prob.Cm)), _p(psd), _p(bisp))
    out["psd_r"], out["bisp_r"] = psd.copy(), None if bisp is None else bisp.copy()
    _intg(f, psd, m2, bisp, m3)
    out["m2mr"], out["m3mr"] = m2.copy(), m3.copy()
    om = TWO_PI * f
    psd2 = np.asfortranarray(psd * (om * om)[:, None])
    m2o2 = np.zeros(nm * nm)
    _intg(f, psd2, m2o2, None, None)
    out["m2o2mr"] = m2o2
    out["modes"] = modes
    return out


def _intg(f, psd, m2, bisp, m3, pulsation=True):
    """intgSpectraVect_ (BsaClassicImpl.F90:567-633)."""
    nf = f.size
    lib().orc_intg_vect(C.c_int(nf), _p(f), C.c_int(1 if pulsation else 0), _p(psd),
                        C.c_int(0 if m2 is None else m2.size), _p(m2), _p(bisp),
                        C.c_int(0 if m3 is None else m3.size), _p(m3))


def intg_weights_2d(nf: int, d: float, rule: str) -> np.ndarray:
    """Per-point weights equivalent to the reference quadratures:
    'vect' intgSpectraVect_ (corner 1-3/2 = -1/2, edge 1/2), 'riemann'
    (Classic scalar, BsaClassicImpl.F90:331,341), 'trapezoid' (zone rule,
    _shared_poly2d.fi:827-832)."""
    w = np.ones((nf, nf))
    if rule == "vect":
        w[0, :] = w[-1, :] = w[:, 0] = w[:, -1] = 0.5
        for a in (0, -1):
            for b in (0, -1):
                w[a, b] = 1.0 - 1.5
    elif rule == "trapezoid":
        e = np.ones(nf)
        e[0] = e[-1] = 0.5
        w = e[:, None] * e[None, :]
    elif rule != "riemann":
        raise ValueError(rule)
    return w * d * d


# ---------------------------------------------------------------------------
# factorised FP64 oracle over point sets
# ---------------------------------------------------------------------------
def a_table(prob, freqs, modes=None) -> np.ndarray:
    """A(n,i;f) = sum_j Psi_a(n,j) S_ji(f), S_ji(f) = rho_ji^|f| sqrt(S_j S_i),
    the projection inside getFM_full_tnm_vect_cls_ (functionsImpl.F90:1450-1476).
    Returns (nm, nnl, ntc, nft)."""
    from bsalib_amd.synth import corr_index
    if modes is None:
        modes = kept_modes(prob)
    P = psi(prob, modes)
    nm, nnl = P.shape[0], P.shape[1]
    freqs = np.asarray(freqs, dtype=np.float64)
    out = np.zeros((nm, nnl, prob.ntc, freqs.size), order="F")
    nl = prob.n_load
    idx = corr_index(nl[:, None], nl[None, :], prob.nn) - 1
    for c, t in enumerate(prob.tc):
        rho = prob.nod_corr[idx, t - 1]
        S = eval_psd(prob, freqs, int(t))                      # (nft, nnl)
        s = np.sqrt(S)
        with np.errstate(divide="ignore"):
            lr = np.log(rho)
        for q, fq in enumerate(freqs):
            af = abs(fq)
            if af == 0.0:
                coh = np.ones_like(rho)        # 0**0 == 1 (Fortran semantics)
            else:
                coh = np.exp(af * lr)
            Sm = coh * (s[q][:, None] * s[q][None, :])
            out[:, :, c, q] = P[:, :, t - 1] @ Sm
    return out


def classic_factorised(prob, rule: str = "vect", truncate: bool = True, nthreads: int = 0,
                       want_bisp: bool = False, point_sample=None) -> dict:
    """Factorised FP64 Classic BFM/BRM + m3 on the uniform grid.  ``truncate``
    reproduces the Classic zero S(f1+f2) beyond +-fmax (functionsImpl.F90:
    1409-1413); otherwise S(f1+f2) is evaluated exactly (Mesher / BsaCL)."""
    L = lib()
    f = prob.freqs()
    nf = f.size
    c = (nf - 1) // 2
    modes = kept_modes(prob)
    nm = len(modes)
    P = psi(prob, modes)
    tcs = prob.tc - 1
    pa = np.asfortranarray(P[:, :, tcs])
    pb = np.asfortranarray(P[:, :, tcs + 3])
    fext = np.arange(-2 * c, 2 * c + 1) * prob.df
    need_ext = not truncate
    ft = fext if need_ext else f
    off = 2 * c if need_ext else c
    A = a_table(prob, ft, modes)
    H = frf(prob, ft, modes)                                   # (nft, nm)
    Hc = np.ascontiguousarray(np.stack([H.real, H.imag], axis=-1))   # (nft, nm, 2)
    d = prob.df * TWO_PI
    W = intg_weights_2d(nf, d, rule)
    if point_sample is None:
        ii, jj = np.meshgrid(np.arange(nf), np.arange(nf), indexing="ij")
        ii, jj = ii.ravel(order="F"), jj.ravel(order="F")
    else:
        ii, jj = point_sample
    k3 = ii + jj - c                                            # grid index of f1+f2
    i1 = (ii - c + off).astype(np.int32)
    i2 = (jj - c + off).astype(np.int32)
    if need_ext:
        i3 = (k3 - c + off).astype(np.int32)
    else:
        i3 = np.where((k3 >= 0) & (k3 < nf), k3, -1).astype(np.int32)
    w = np.ascontiguousarray(W[ii, jj])
    f3 = f[ii] + f[jj]
    H3 = frf(prob, f3, modes)                                  # (npts, nm)
    H3c = np.ascontiguousarray(np.stack([H3.real, H3.imag], axis=-1))
    m3f = np.zeros(nm ** 3)
    m3r = np.zeros(nm ** 3)
    npts = ii.size
    bfm = np.zeros((nm ** 3, npts), order="F") if want_bisp else None
    brm = np.zeros((nm ** 3, npts), order="F") if want_bisp else None
    L.orc_bfm_factorised(C.c_int(npts), _p(i1, C.c_int), _p(i2, C.c_int), _p(i3, C.c_int),
                         _p(w), C.c_int(nm), C.c_int(prob.nnl), C.c_int(prob.ntc), _p(pa),
                         _p(pb), _p(np.asfortranarray(A)), C.c_int(ft.size),
                         _p(Hc), _p(H3c),
                         _p(m3f), _p(m3r), _p(bfm), _p(brm), C.c_int(nthreads))
    return {"m3mf": m3f, "m3mr": m3r, "bfm": bfm, "brm": brm, "modes": modes,
            "points": (ii, jj)}


def modal_psd(prob, rule_pulsation: bool = True) -> dict:
    """Modal force/response PSD and m2 on the Classic grid:
    psd(n,p;f) = sum_jk Psi_a(n,j) Psi_a(p,k) S_jk(f) (functionsImpl.F90:1546-1558),
    psd_r = psd * Re[H_p conj(H_n)] (:1686-1689); m2 by intgSpectraVect_."""
    f = prob.freqs()
    modes = kept_modes(prob)
    nm = len(modes)
    P = psi(prob, modes)
    A = a_table(prob, f, modes)                                 # (nm, nnl, ntc, nf)
    psd = np.zeros((f.size, nm * nm), order="F")
    for c, t in enumerate(prob.tc):
        # psd(q, n + p*nm) = sum_k A(n,k) Psi_a(p,k)
        S = np.einsum("nkq,pk->qnp", A[:, :, c, :], P[:, :, t - 1])
        psd += S.reshape(f.size, nm * nm, order="F")
    H = frf(prob, f, modes)
    fr = (H[:, None, :] * np.conj(H[:, :, None])).real        # [q, n, p] = Re[H_p conj(H_n)]
    psd_r = psd * fr.reshape(f.size, nm * nm, order="F")
    om = TWO_PI * f
    d = prob.df * TWO_PI
    wv = np.full(f.size, d)
    wv[0] = wv[-1] = d / 2.0
    return {"psd_f": psd, "psd_r": psd_r, "m2mf": wv @ psd, "m2mr": wv @ psd_r,
            "m2o2mr": wv @ (psd_r * (om * om)[:, None])}


# ---------------------------------------------------------------------------
# Mesher per-point BFM (non-POD) -- literal restatement
# ---------------------------------------------------------------------------
def msh_points(prob, fi, fj) -> np.ndarray:
    """getFM_full_tnm_scalar_msh_ (functionsImpl.F90:247-465) at each point;
    returns bfm(nm^3, npts).  PSDs at fi, fj, fi+fj via evalPSD (:302-306)."""
    L = lib()
    modes = kept_modes(prob)
    nm = len(modes)
    fi = np.atleast_1d(np.asarray(fi, dtype=np.float64))
    fj = np.atleast_1d(np.asarray(fj, dtype=np.float64))
    out = np.zeros((nm ** 3, fi.size), order="F")
    buf = np.zeros(nm ** 3)
    phi = _f64F(prob.phi)
    wfc = _f64F(prob.wfc)
    corr = _f64F(prob.nod_corr)
    for q in range(fi.size):
        S1 = np.asfortranarray(np.stack([eval_psd(prob, [fi[q]], int(t))[0] for t in prob.tc], axis=1))
        S2 = np.asfortranarray(np.stack([eval_psd(prob, [fj[q]], int(t))[0] for t in prob.tc], axis=1))
        S3 = np.asfortranarray(np.stack([eval_psd(prob, [fi[q] + fj[q]], int(t))[0] for t in prob.tc], axis=1))
        L.orc_bfm_point_msh(C.c_double(fi[q]), C.c_double(fj[q]), C.c_int(prob.nnl), C.c_int(prob.nn),
                            C.c_int(prob.nlibs), C.c_int(prob.libs_load.size),
                            _p(_i32(prob.libs_load), C.c_int), C.c_int(nm), _p(modes, C.c_int),
                            C.c_int(prob.phi.shape[1]), C.c_int(prob.ntc), _p(_i32(prob.tc), C.c_int),
                            _p(_i32(prob.n_load), C.c_int), _p(wfc), C.c_int(wfc.shape[1]), _p(phi),
                            C.c_int(phi.shape[0]), _p(corr), C.c_long(corr.shape[0]), _p(S1),
                            _p(S2), _p(S3), _p(buf))
        out[:, q] = buf
    return out


# ---------------------------------------------------------------------------
# Mesher per-point BFM through the POD of S(f) -- literal restatement
# ---------------------------------------------------------------------------
def pod_threshold_count(vals_desc, rlim: float) -> int:
    """getNPODModesByThreshold_ (functionsImpl.F90:203-230) on values in
    decreasing order, as its dgesvd build reads them (__IDX = 1, __OP = +):
    1 when all values are equal, else the leading run of values >= rlim *
    largest.  The loop is bounded by the number of values (the reference's is
    not)."""
    v = np.asarray(vals_desc, dtype=np.float64)
    if np.all(v == v[0]):
        return 1
    limval = rlim * v[0]
    k = 2                                   # 1-based, as the reference
    while k <= v.size and v[k - 1] >= limval:
        k += 1
    return k - 1


def nodal_psd_matrix(prob, f: float, tc: int) -> np.ndarray:
    """getFullNodalPSD (WindPSDImpl.F90:76-108) with idir = 1, as the POD BFM
    calls it (functionsImpl.F90:628-629, 770): PSDmat(j, i) = sqrt(S_j S_i) *
    nod_corr(id(j, i), 1)**|f|."""
    from bsalib_amd.synth import corr_index
    nl = prob.n_load
    idx = corr_index(nl[:, None], nl[None, :], prob.nn) - 1
    S = eval_psd(prob, [f], int(tc))[0]
    return np.sqrt(S[:, None] * S[None, :]) * prob.nod_corr[idx, 0] ** abs(f)


def msh_points_pod(prob, fi, fj, nmodes: int = 0, lim: float = 0.0) -> np.ndarray:
    """getFM_full_tm_scalar_msh_POD_ (functionsImpl.F90:474-1019; dsyev build,
    no POD data caching) at each point: S(fi), S(fj), S(fi+fj) decomposed
    (eigenvalues ascending, as dsyev), the kept pairs taken from the largest
    down (:840-997), then blocks 5-2-2 (:858-899), 2-2-5 (:902-941) and 2-5-2
    (:945-997).  Kept pairs (:803-823): lim > 0 -> pod_threshold_count per
    matrix; else nmodes > 0 -> nmodes; else all.  Returns bfm(nm^3, npts)."""
    modes = kept_modes(prob)
    P = psi(prob, modes)
    nm, n = len(modes), prob.nnl
    fi = np.atleast_1d(np.asarray(fi, dtype=np.float64))
    fj = np.atleast_1d(np.asarray(fj, dtype=np.float64))
    out = np.zeros((nm ** 3, fi.size), order="F")

    def decomp(f, t):
        lam, V = np.linalg.eigh(nodal_psd_matrix(prob, f, t))
        if lim > 0.0:
            k = pod_threshold_count(lam[::-1], lim)
        elif nmodes > 0:
            k = min(nmodes, n)
        else:
            k = n
        keep = np.arange(n - 1, n - 1 - k, -1)          # NNODESL down to NNODESL-(k-1)
        return lam[keep], V[:, keep]

    for q in range(fi.size):
        B = np.zeros((nm, nm, nm))                     # B[m, n, o] = bfm(posm)
        for t in prob.tc:
            Pl, Pq = P[:, :, t - 1], P[:, :, t + 2]    # tc, tcP3 = tc + 3 (:561-562)
            D1, V1 = decomp(fi[q], t)
            D2, V2 = decomp(fj[q], t)
            D3, V3 = decomp(fi[q] + fj[q], t)
            L2, L3 = Pl @ V2, Pl @ V3                  # Z_lin_w2, Z_lin_w1w2
            for p in range(D1.size):
                vp = V1[:, p]
                t1 = Pl @ vp                           # V_lin_w1
                # 5-2-2: 2 VZ_quad_w1w2(m) V_lin_w1(n) Z_lin_w2(o) D_p T_q
                Q = Pq @ (vp[:, None] * V2)
                B += 2.0 * D1[p] * np.einsum("mq,n,oq->mno", Q * D2, t1, L2)
                # 2-2-5: 2 Z_lin_w1w2(m) V_lin_w1(n) VZ_quad_w1w1w2(o) D_p T_q
                Q = Pq @ (vp[:, None] * V3)
                B += 2.0 * D1[p] * np.einsum("mq,n,oq->mno", L3 * D3, t1, Q)
            for p in range(D3.size):
                vp = V3[:, p]
                t1 = Pl @ vp                           # V_lin_w1w2
                # 2-5-2: 2 V_lin_w1w2(m) VZ_quad_w1w2w2(n) Z_lin_w2(o) D_p T_q
                Q = Pq @ (vp[:, None] * V2)
                B += 2.0 * D3[p] * np.einsum("m,nq,oq->mno", t1, Q * D2, L2)
        out[:, q] = B.reshape(-1, order="F")
    return out


def bsacl_run(prob, fi, fj, psd_id=None, nthreads: int = 0) -> np.ndarray:
    """bsaclRun with bfm_kernel id 2 (BsaCL/src/bsacl.c:1307-1592,
    BsaCL/src/bsacl.cl:405-930) in FP64: response third moment over all
    (fi, fj) pairs, plain Riemann sum times dInfl."""
    L = lib()
    modes = kept_modes(prob)
    nm = len(modes)
    P = psi(prob, modes)
    mo = modes - 1
    Mg = np.ascontiguousarray(prob.Mm[mo])
    Kg = np.ascontiguousarray(prob.Km[mo])
    Cg = np.ascontiguousarray(prob.Cm[np.ix_(mo, mo)].T)        # Cg[mi*nm+mi] diagonal
    fi = np.ascontiguousarray(fi, dtype=np.float64)
    fj = np.ascontiguousarray(fj, dtype=np.float64)
    res = np.zeros(nm ** 3)
    L.orc_bsacl_run(C.c_int(psd_id or prob.psd_id), C.c_int(prob.ntc), _p(_i32(prob.tc), C.c_int),
                    C.c_int(prob.nnl), C.c_int(nm), _p(_i32(prob.n_load), C.c_int), _p(P),
                    C.c_int(prob.nn), _p(_f64F(prob.nod_corr)), _p(_f64F(prob.u_node)),
                    _p(_f64F(prob.turb_scales)), _p(_f64F(prob.turb_std)), _p(Mg), _p(Cg), _p(Kg),
                    C.c_int(fi.size), _p(fi), C.c_int(fj.size), _p(fj), C.c_int(1), _p(res),
                    C.c_int(nthreads))
    return res


# ---------------------------------------------------------------------------
# the reference itself (oracle/_ref/ref_harness, built by oracle/ref/build_ref.sh)
# ---------------------------------------------------------------------------
REF_HARNESS = HERE / "_ref" / "ref_harness"


def reference_available() -> bool:
    return REF_HARNESS.exists()


def _write_ref_input(prob, path, mode: int, fi=None, fj=None):
    modes = kept_modes(prob)
    P = psi(prob, modes)
    npts = 0 if fi is None else int(np.size(fi))
    hdr = np.array([prob.nn, prob.nlibs, prob.nnl, prob.libs_load.size, prob.phi.shape[1],
                    modes.size, prob.ntc, prob.wfc.shape[1], prob.nwz, prob.psd_id,
                    prob.nfreqs, mode, npts, prob.nod_corr.shape[0]], dtype=np.int32)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        for a in (prob.n_load, prob.libs_load, modes, prob.tc, prob.wz_node):
            fh.write(np.asarray(a, dtype=np.int32).tobytes())
        fh.write(np.float64(prob.df).tobytes())
        for a in (prob.phi, prob.natf, prob.Mm, prob.Km, prob.Cm, prob.u_node, prob.turb_scales,
                  prob.turb_std, prob.u_mean_ref, prob.nod_corr, prob.wfc, P):
            fh.write(np.asarray(a, dtype=np.float64).tobytes(order="F"))
        if npts:
            fh.write(np.asarray(fi, dtype=np.float64).tobytes())
            fh.write(np.asarray(fj, dtype=np.float64).tobytes())


def reference_run(prob, mode: int = 1, fi=None, fj=None, workdir=None,
                  malloc_perturb: int = 255) -> dict:
    """Run the reference's own operators through oracle/_ref/ref_harness.
    mode bit 1: Classic vect BFM + BRM on the full grid; bit 2: Mesher BFM at (fi, fj);
    bit 4: the ONLY_DIAG Classic model (getFM_diag_tnlm_vect_cls_ / getRM_diag_vect_cls_).

    The Classic BFM reads S(f1+f2) beyond +-fmax from padding it never
    initialises (functionsImpl.F90:1409-1413); glibc's MALLOC_PERTURB_=255
    makes fresh allocations zero-filled, which is the intended zero padding.
    Any other value exposes which outputs depend on that padding."""
    import tempfile
    if not reference_available():
        raise FileNotFoundError(str(REF_HARNESS))
    modes = kept_modes(prob)
    nm = modes.size
    nf = prob.nf
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        _write_ref_input(prob, fin, mode, fi, fj)
        subprocess.run([str(REF_HARNESS), fin, fout], check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"),
                                MALLOC_PERTURB_=str(malloc_perturb)))
        raw = np.fromfile(fout, dtype=np.float64)
    out = {}
    o = 0

    def take(n, shape):
        nonlocal o
        a = raw[o:o + n].reshape(shape, order="F")
        o += n
        return a
    out["suvw"] = take(nf * prob.nnl * prob.ntc, (nf, prob.nnl * prob.ntc))
    if mode & 1:
        out["psd_f"] = take(nf * nm * nm, (nf, nm * nm))
        out["bisp_f"] = take(nf * nf * nm ** 3, (nf, nf, nm ** 3))
        out["psd_r"] = take(nf * nm * nm, (nf, nm * nm))
        out["bisp_r"] = take(nf * nf * nm ** 3, (nf, nf, nm ** 3))
    if mode & 4:
        out["psd_f_diag"] = take(nf * nm, (nf, nm))
        out["bisp_f_diag"] = take(nf * nf * nm, (nf, nf, nm))
        out["psd_r_diag"] = take(nf * nm, (nf, nm))
        out["bisp_r_diag"] = take(nf * nf * nm, (nf, nf, nm))
    if mode & 2 and fi is not None:
        out["bfm_msh"] = take(nm ** 3 * np.size(fi), (nm ** 3, np.size(fi)))
    return out


def reference_time(prob, threads: int) -> float:
    """Wall time of the reference's getFM_full_tnm_vect_cls_ on ``prob``."""
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        _write_ref_input(prob, fin, 1)
        r = subprocess.run([str(REF_HARNESS), fin, os.path.join(td, "o.bin"), "time"],
                           check=True, capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
    for line in r.stdout.splitlines():
        if line.startswith("REFTIME"):
            return float(line.split()[1])
    raise RuntimeError("reference timing line missing:\n" + r.stdout[-2000:])


def mesher_integrate(prob, geom, levels, modes=None, points=None, brm_out=None, diag=False):
    """Zone quadrature of the Mesher on given tiles (CPU restatement of
    bsahip_mesher_run): per tile, the force bispectrum at the level-l grid by
    getFM_full_tnm_scalar_msh_ (msh_points), bilinear interpolation onto the
    BRM grid, the FRF product Re[H_m(f1) H_n(f2) conj(H_p(f1+f2))] and
    trapezoid weights (zones/_shared_poly2d.fi:827-832: vertex 1/4, edge 1/2,
    interior 1, times dw1 dw2).  geom rows: f1a f1b f2a f2b n1 n2 pol1 pol2
    sub1 sub2 (sub = BRM cells per level-0 cell).  Returns m3mf, m3mr and the
    BFM evaluation counts: level-0 points of all tiles counted once (they lie on
    one tensor grid), refinement points per tile.  points: the point BFM
    (default msh_points; e.g. a msh_points_pod partial for the POD BFM).
    brm_out: a list that receives, per tile, the BRM grid frequencies f1
    (NA+1), f2 (NB+1) and the integrand without weights, brm (NM^3, NA+1, NB+1).
    diag: the ONLY_DIAG model -- NM entries per point (default points
    msh_points_diag) and the FRF product Re[H_m(f1) H_m(f2) conj(H_m(f1+f2))]."""
    modes = kept_modes(prob) if modes is None else np.asarray(modes)
    nm = len(modes)
    nm3 = nm if diag else nm ** 3
    if points is None:
        points = msh_points_diag if diag else msh_points
    m3f = np.zeros(nm3)
    m3r = np.zeros(nm3)
    ev1 = 0
    for t, lv in zip(np.asarray(geom), np.broadcast_to(np.asarray(levels), (len(geom),))):
        a1, b1, a2, b2 = t[:4]
        n1, n2 = int(t[4]), int(t[5])
        sub1, sub2 = int(t[8]), int(t[9])
        lv = int(lv)
        nb1, nb2 = n1 * lv + 1, n2 * lv + 1
        g1 = a1 + np.arange(nb1) * (b1 - a1) / (nb1 - 1)
        g2 = a2 + np.arange(nb2) * (b2 - a2) / (nb2 - 1)
        g1[-1], g2[-1] = b1, b2
        F1, F2 = np.meshgrid(g1, g2, indexing="ij")
        B = points(prob, F1.ravel(), F2.ravel()).reshape(nm3, nb1, nb2)
        if lv > 1:
            ev1 += nb1 * nb2 - (n1 + 1) * (n2 + 1)
        s1, s2 = max(1, sub1 // lv), max(1, sub2 // lv)
        NA, NB = (nb1 - 1) * s1, (nb2 - 1) * s2
        h1, h2 = (b1 - a1) / NA, (b2 - a2) / NB

        def interp(nb, s, N):
            P = np.zeros((N + 1, nb))
            for a in range(N + 1):
                c = min(a // s, nb - 2)
                u = (a - c * s) / s
                P[a, c] += 1.0 - u
                P[a, c + 1] += u
            return P
        P1, P2 = interp(nb1, s1, NA), interp(nb2, s2, NB)
        Bb = np.einsum("ai,xij,bj->xab", P1, B, P2)                    # (nm3, NA+1, NB+1)
        f1 = a1 + np.arange(NA + 1) * h1
        f2 = a2 + np.arange(NB + 1) * h2
        w1 = np.ones(NA + 1); w1[[0, -1]] = 0.5
        w2 = np.ones(NB + 1); w2[[0, -1]] = 0.5
        W = np.outer(w1, w2) * (TWO_PI * h1) * (TWO_PI * h2)
        H1 = frf(prob, f1, modes)                                       # (NA+1, nm)
        H2 = frf(prob, f2, modes)                                       # (NB+1, nm)
        H3 = frf(prob, f1[:, None] + f2[None, :], modes)               # (NA+1, NB+1, nm)
        if diag:
            R = np.einsum("am,bm,abm->mab", H1, H2, np.conj(H3)).real
        else:
            R = np.einsum("am,bn,abp->mnpab", H1, H2, np.conj(H3)).real.reshape(
                nm3, NA + 1, NB + 1, order="F")
        m3f += np.einsum("xab,ab->x", Bb, W)
        m3r += np.einsum("xab,xab,ab->x", Bb, R, W)
        if brm_out is not None:
            brm_out.append({"f1": f1, "f2": f2, "brm": Bb * R})
    g = np.asarray(geom)
    nodes = set()
    for t in g[g[:, 2] == g[0, 2]]:
        n = int(t[4])
        nodes.update(np.round(t[0] + np.arange(n + 1) * (t[1] - t[0]) / n, 12).tolist())
    ev0 = len(nodes) ** 2
    return {"m3mf": m3f, "m3mr": m3r, "evals_pre": ev0, "evals_post": ev1}


# ---------------------------------------------------------------------------
# post-processing: BR decomposition, modal recombination, peak factors
# ---------------------------------------------------------------------------
def br_sfm(prob, fnat: float, m: int) -> float:
    """getBR_SFm_val_ (functionsImpl.F90:2650-2743): modal force PSD of mode m
    (1-based, full list) at fnat, summed over components and loaded node pairs
    with their libs (coherence column tc, PSDs by evalPSD at |fnat|)."""
    from bsalib_amd.synth import corr_index
    nl = prob.n_load
    fi = abs(fnat)
    idx = corr_index(nl[:, None], nl[None, :], prob.nn) - 1      # [j, k]
    psd = 0.0
    for t in prob.tc:
        S = eval_psd(prob, [fnat], int(t))[0]                   # (nnl,)
        rho = prob.nod_corr[idx, t - 1]
        SJK = rho ** fi * np.sqrt(S[:, None] * S[None, :])      # [j, k]
        rows = (nl - 1)[:, None] * prob.nlibs + prob.libs_load[None, :] - 1   # [node, lib]
        ph = prob.phi[rows, m - 1]                              # phi_(pos_n + libs, m)
        a = prob.wfc[:, t - 1, :].T                             # wfc(libs, tc, node) -> [node, lib]
        # sum(matmul(phij, phik) * matmul(akU, ajU) * S_JK): element (a, b) pairs
        # phij(a) with akU(a) and phik(b) with ajU(b) (:2717-2727)
        pj_ak = ph[:, None, :] * a[None, :, :]                  # [j, k, lib]: phij(a) * akU(a)
        pk_aj = ph[None, :, :] * a[:, None, :]                  # [j, k, lib]: phik(b) * ajU(b)
        psd += float((pj_ak.sum(-1) * pk_aj.sum(-1) * SJK).sum())
    return psd


def br_decomp(prob, m2mf, modes=None):
    """bsa_computeBRdecomp (BsaLibImpl.F90:1292-1421): background m2mf(m,m)/Km^2
    and resonant pi * 2pi * f_n * S_Fm(f_n) / (2 xsi Km^2) per kept mode."""
    modes = kept_modes(prob) if modes is None else np.asarray(modes)
    nm = len(modes)
    xsi = np.asarray(prob.meta["xi"], dtype=np.float64)
    bkg, res = np.zeros(nm), np.zeros(nm)
    for im, m in enumerate(modes):
        k2 = prob.Km[m - 1] ** 2
        bkg[im] = m2mf[im * nm + im] / k2
        f = prob.natf[m - 1]
        res[im] = np.pi * TWO_PI * f * br_sfm(prob, f, int(m)) / (2 * xsi[m - 1] * k2)
    return bkg, res


def modal_recombination(phi, m2mr, m3mr, m2o2mr=None) -> dict:
    """modalRecombination (app/bsa.F90:935-1053): unilateral (diag) and complete
    cubic combination (full) nodal m2, m3, m2o2 and sk = m3 / m2^(3/2)."""
    phi = np.asarray(phi, dtype=np.float64)
    ndofs, nm = phi.shape
    m2 = np.asarray(m2mr, dtype=np.float64).reshape(nm, nm, order="F")
    m3 = np.asarray(m3mr, dtype=np.float64).reshape(nm, nm, nm, order="F")
    m22 = None if m2o2mr is None else np.asarray(m2o2mr, dtype=np.float64).reshape(nm, nm, order="F")
    o = {k: np.zeros(ndofs) for k in ("m2_diag", "m3_diag", "m2o2_diag",
                                      "m2_full", "m3_full", "m2o2_full")}
    for im in range(nm):                                    # :959-984
        r = phi[:, im] * phi[:, im]
        o["m2_diag"] += m2[im, im] * r
        if m22 is not None:
            o["m2o2_diag"] += m22[im, im] * r
        r = r * phi[:, im]
        o["m3_diag"] += m3[im, im, im] * r
    for k in range(nm):                                     # :1003-1043
        p3 = phi[:, k]
        for j in range(nm):
            p2 = phi[:, j]
            o["m2_full"] += m2[j, k] * p2 * p3
            if m22 is not None:
                o["m2o2_full"] += m22[j, k] * p2 * p3
            for i in range(nm):
                o["m3_full"] += m3[i, j, k] * phi[:, i] * p2 * p3
    o["sk_diag"] = o["m3_diag"] / o["m2_diag"] ** 1.5
    o["sk_full"] = o["m3_full"] / o["m2_full"] ** 1.5
    return o


def peak_factors(m2, m2o2, obs_time: float, sk=None):
    """bsa_computePeakFactors (BsaLibImpl.F90:1429-1519): Gaussian peak factor
    beta + gamma/beta with beta = sqrt(2 ln(sqrt(m2o2/m2) T / 2pi)), and with a
    skewness the Hermite-model (Kwon & Kareem) non-Gaussian peak factors.
    Returns (peak_g, peak_ng_pos, peak_ng_neg), or None where the reference
    returns early (m2o2 all zero, or some beta < 1)."""
    gamma = 0.5772
    m2 = np.asarray(m2, dtype=np.float64)
    m2o2 = np.asarray(m2o2, dtype=np.float64)
    if np.all(m2o2 == 0):
        return None
    beta = np.sqrt(m2o2 / m2) / TWO_PI * obs_time
    if np.any(beta < 1.0):
        return None
    beta = np.sqrt(2 * np.log(beta))
    peak_g = gamma / beta + beta
    if sk is None:
        return peak_g, None, None
    sk = np.asarray(sk, dtype=np.float64)
    g4 = 1.5 * sk * sk
    h40 = np.sqrt(1.0 + 1.5 * g4)
    h3 = sk / (4.0 + 2.0 * h40)
    h4 = (h40 - 1.0) / 18.0
    k = 1.0 / np.sqrt(1 + 2.0 * h3 * h3 + 6.0 * h4 * h4)
    b2, b3 = beta * beta, beta * beta * beta
    pos = 5.44 / b3
    pos = pos + 3.0 / beta * (np.pi ** 2 / 6.0 - gamma + gamma ** 2)
    pos = pos + (b3 + 3 * beta * (gamma - 1))
    pos = pos * h4
    neg = pos.copy()
    r = h3 * (b2 + 2.0 * gamma - 1.0 + 1.98 / b2)
    pos, neg = pos + r, neg - r
    return peak_g, (pos + peak_g) * k, (neg + peak_g) * k


# ---------------------------------------------------------------------------
# ONLY_DIAG Classic model -- numpy restatement
# ---------------------------------------------------------------------------
def msh_points_diag(prob, fi, fj) -> np.ndarray:
    """getFM_diag_tnm_scalar_msh_ (functionsImpl.F90:1123-1216) at each point,
    returns bfm(nm, npts): per loaded node n and component,
      2 Psi_q Psi_l^2 [S_n(f1) S_n(f2) + S_n(f1+f2) S_n(f2) + S_n(f1+f2) S_n(f1)]
    (after the lib sums, as classic_diag), S(f1+f2) evaluated at the point.
    The reference forms the second and third products as S(f1) S(f2)^2 and
    S(f1)^2 S(f2) (Suvw_IJJ, Suvw_IJI at :1161-1163) and leaves S(f1+f2),
    which it evaluates, unused; this follows the Classic model instead
    (classic_diag), so that both approaches integrate one spectrum."""
    modes = kept_modes(prob)
    P = psi(prob, modes)
    fi = np.atleast_1d(np.asarray(fi, dtype=np.float64))
    fj = np.atleast_1d(np.asarray(fj, dtype=np.float64))
    out = np.zeros((len(modes), fi.size))
    for t in prob.tc:
        S1 = eval_psd(prob, fi, int(t))                 # (npts, nnl)
        S2 = eval_psd(prob, fj, int(t))
        S3 = eval_psd(prob, fi + fj, int(t))
        W = P[:, :, t + 2] * P[:, :, t - 1] ** 2        # (nm, nnl)
        out += 2.0 * W @ (S1 * S2 + S3 * S2 + S3 * S1).T
    return out


def classic_diag(prob) -> dict:
    """getFM_diag_tnlm_vect_cls_ (functionsImpl.F90:1747-1970) and
    getRM_diag_vect_cls_ (:1975-2101) on the Classic grid, then
    intgSpectraVect_: each loaded node on its own (no coherence between nodes),
    diagonal modal entries only.  Per mode m, with Psi_l / Psi_q the linear /
    quadratic projections of node n and S_n the nodal PSD:
      psd(f)      = sum_{tc,n} Psi_l^2 S_n(f)
      bisp(f1,f2) = 2 sum_{tc,n} Psi_q Psi_l^2 [S_n(f1) S_n(f2)
                      + S_n(f1+f2) S_n(f2) + S_n(f1+f2) S_n(f1)]
    (the three lib products of :1899-1903 collapse to Psi_q Psi_l^2 after the
    lib sums), S(f1+f2) = 0 beyond +-fmax (zero padding, T2); response spectra
    times |H_m(f)|^2 and Re[H_m(f1) H_m(f2) conj H_m(f1+f2)], H at w1 + w2.
    Returns psd_f, psd_r (nf, nm), bisp_f, bisp_r (nf, nf, nm) and m2mf,
    m2mr, m3mf, m3mr (nm)."""
    modes = kept_modes(prob)
    P = psi(prob, modes)
    nm = modes.size
    f = prob.freqs()
    nf = f.size
    c = nf // 2
    psd = np.zeros((nf, nm))
    bisp = np.zeros((nf, nf, nm))
    ii, jj = np.meshgrid(np.arange(nf), np.arange(nf), indexing="ij")
    s = ii + jj - c                                     # grid index of f1 + f2
    inside = (s >= 0) & (s < nf)
    for t in prob.tc:
        S = eval_psd(prob, f, int(t))                   # (nf, nnl)
        Pl, Pq = P[:, :, t - 1], P[:, :, t + 2]
        psd += S @ (Pl ** 2).T
        for n in range(prob.nnl):
            Sn = S[:, n]
            Ss = np.where(inside, Sn[np.clip(s, 0, nf - 1)], 0.0)
            tmp = Sn[:, None] * Sn[None, :] + Ss * Sn[None, :] + Ss * Sn[:, None]
            bisp += 2.0 * tmp[:, :, None] * (Pq[:, n] * Pl[:, n] ** 2)[None, None, :]
    mo = modes - 1
    w = TWO_PI * f
    H = 1.0 / (prob.Km[mo][None, :] - (w * w)[:, None] * prob.Mm[mo][None, :]
               + 1j * w[:, None] * np.diag(prob.Cm)[mo][None, :])            # (nf, nm)
    w12 = w[:, None] + w[None, :]
    H12 = 1.0 / (prob.Km[mo][None, None, :] - (w12 * w12)[:, :, None] * prob.Mm[mo][None, None, :]
                 + 1j * w12[:, :, None] * np.diag(prob.Cm)[mo][None, None, :])
    psd_r = psd * np.abs(H) ** 2
    bisp_r = bisp * (H[:, None, :] * H[None, :, :] * np.conj(H12)).real
    out = dict(psd_f=psd, psd_r=psd_r, bisp_f=bisp, bisp_r=bisp_r)
    for k, (ps, bs) in dict(f=(psd, bisp), r=(psd_r, bisp_r)).items():
        m2, m3 = np.zeros(nm), np.zeros(nm)
        _intg(f, np.asfortranarray(ps), m2, np.asfortranarray(bs), m3)
        out[f"m2m{k}"], out[f"m3m{k}"] = m2, m3
    return out
