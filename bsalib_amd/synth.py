"""Seeded synthetic structures and wind fields for the BFM hot path.

The reference takes every one of these arrays as user data through its
setters (``bsa_setTotalNOfNodes``, ``bsa_setModalInfo``, ``bsa_setWindFCoeffs``,
``bsa_setSpatialNodalCorr`` ... in src/BsaLib.F90, canonical call order in
app/bsa.F90:525-606).  This module builds them, in the reference's own array
shapes and Fortran (column-major) layouts, for two geometries:

* ``tower``  vertical cantilever, nodes at z_i = i*H/NN, along-wind DOF;
* ``deck``   horizontal bridge deck at constant height, vertical DOF.

Conventions that matter for parity with the reference:

* mode shapes are divided by their max |phi| so that max|phi| == 1 exactly:
  the reference silently drops any other mode (src/bsa/BsaLibImpl.F90:545-599);
* the nodal coherence is stored packed, upper triangle over *total* node ids,
  one column per turbulence component (``nod_corr((NN*NN+NN)/2, 3)``,
  index rule src/utils/utility.F90:60-79), and holds rho_ij so that the
  coherence at frequency f is rho_ij**|f| (functionsImpl.F90:1450-1476);
* wind force coefficients ``wfc(NLIBSL, NDEGW=6, NNL)``: column ``tc`` holds the
  linear (U*u) coefficient, column ``tc+3`` the quadratic (u*u) one
  (src/bsa/BsaLibImpl.F90:643-663, functionsImpl.F90:1438-1439).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

RHO_AIR = 1.22


@dataclass
class Problem:
    """One structure + wind field + Classic grid, in the reference's layouts."""

    nn: int                      # total nodes
    nlibs: int                   # DOFs per node
    n_load: np.ndarray           # (NNL,) int32, 1-based loaded node ids
    libs_load: np.ndarray        # (NLIBSL,) int32, 1-based loaded DOFs
    phi: np.ndarray              # (NDOFS, NM) float64, Fortran order
    natf: np.ndarray             # (NM,) natural frequencies [Hz]
    Mm: np.ndarray               # (NM,) modal masses
    Km: np.ndarray               # (NM,) modal stiffnesses
    Cm: np.ndarray               # (NM, NM) modal damping matrix (diagonal used)
    psd_id: int                  # 1 von Karman, 2 Kaimal, 3 Davenport-Greisch, 5 Davenport-Uliege
    tc: np.ndarray               # (NTC,) int32, 1-based turbulence components (1=u,2=v,3=w)
    wz_node: np.ndarray          # (NN,) int32, 1-based wind zone per node
    u_node: np.ndarray           # (NN,) mean wind speed per node
    turb_scales: np.ndarray      # (3, 3, NWZ) L(component, direction, zone)
    turb_std: np.ndarray         # (3, NWZ) sigma(component, zone)
    u_mean_ref: np.ndarray       # (NWZ,) zone reference mean speed
    nod_corr: np.ndarray         # ((NN*NN+NN)//2, 3) packed rho_ij per component
    wfc: np.ndarray              # (NLIBSL, 6, NNL) wind force coefficients
    nfreqs: int                  # half grid count: NF = 2*(nfreqs-1)+1
    df: float                    # grid step [Hz]
    name: str = "problem"
    coords: Optional[np.ndarray] = None   # (NN,) position along the structure [m]
    meta: dict = field(default_factory=dict)

    # ---- derived sizes -------------------------------------------------
    @property
    def nnl(self) -> int:
        return int(self.n_load.size)

    @property
    def nm(self) -> int:
        return int(self.phi.shape[1])

    @property
    def ntc(self) -> int:
        return int(self.tc.size)

    @property
    def nwz(self) -> int:
        return int(self.u_mean_ref.size)

    @property
    def nf(self) -> int:
        """Classic grid size under the pulsation convention (always odd)."""
        return 2 * (self.nfreqs - 1) + 1

    def freqs(self) -> np.ndarray:
        """Classic grid f_k = k*df, k = -(nfreqs-1)..(nfreqs-1)
        (src/bsa/classic/BsaClassicImpl.F90:534-555)."""
        n1 = self.nfreqs - 1
        return np.arange(-n1, n1 + 1, dtype=np.float64) * self.df


def corr_index(ni: np.ndarray, nj: np.ndarray, tot: int) -> np.ndarray:
    """1-based packed upper-triangle index of (ni, nj) over ``tot`` nodes
    (src/utils/utility.F90:60-79: id = (a-1)*tot + b - (a*a-a)/2 with a<=b)."""
    ni = np.asarray(ni, dtype=np.int64)
    nj = np.asarray(nj, dtype=np.int64)
    a = np.minimum(ni, nj)
    b = np.maximum(ni, nj)
    return (a - 1) * tot + b - (a * a - a) // 2


def pack_corr(rho_full: np.ndarray) -> np.ndarray:
    """Pack a symmetric (NN, NN) matrix into the reference's vector layout."""
    nn = rho_full.shape[0]
    out = np.zeros((nn * nn + nn) // 2, dtype=np.float64)
    ii, jj = np.triu_indices(nn)
    out[corr_index(ii + 1, jj + 1, nn) - 1] = rho_full[ii, jj]
    return out


def _normalise_modes(phi: np.ndarray) -> np.ndarray:
    m = np.abs(phi).max(axis=0)
    out = phi / m[None, :]
    # the entry that realised the max becomes +-1 exactly (x/x == 1 in IEEE)
    assert np.all(np.abs(out).max(axis=0) == 1.0)
    return np.asfortranarray(out)


def _modal_matrices(natf: np.ndarray, xi: np.ndarray):
    om = 2.0 * np.pi * natf
    Mm = np.ones_like(natf)
    Km = om * om * Mm
    Cm = np.diag(2.0 * xi * om * Mm)
    return Mm, Km, np.asfortranarray(Cm)


def _vk_scales(ntc_comp):
    return {1: 100.0, 2: 40.0, 3: 20.0}


def tower(nn: int = 20, nm: int = 3, comps=(1,), natf=None, xi=None, nfreqs: int = 65,
          df: float = 1.0 / 32.0, seed: int = 0, height: float = 100.0, u10: float = 30.0,
          alpha: float = 0.16, psd_id: int = 1, decay: float = 10.0,
          intensities=(0.15, 0.10, 0.08), scales=(100.0, 40.0, 20.0),
          cd: float = 1.2, area: float = 2.0, zone_per_node: bool = True, nlibs: int = 1,
          libs_load=(1,)) -> Problem:
    """Vertical cantilever tower (configuration 1 by default: NN=20, NM=3, u only,
    f_n = 0.2/0.6/1.2 Hz, 1 % damping, 129-point grid over [-2, 2] Hz).
    nlibs > 1: DOFs per node (DOF 1 along-wind, the others rotations about the
    tower axis whose mode shapes are seeded sinusoids); libs_load lists the
    loaded DOFs, a rotation carrying the drag times a lever arm of 0.4 m."""
    rng = np.random.default_rng(seed)
    z = np.arange(1, nn + 1, dtype=np.float64) * height / nn
    U = u10 * (z / 10.0) ** alpha
    comps = tuple(int(c) for c in comps)
    nwz = nn if zone_per_node else 1
    wz_node = (np.arange(1, nn + 1) if zone_per_node else np.ones(nn)).astype(np.int32)
    u_ref = U.copy() if zone_per_node else np.array([u10 * (height / 10.0) ** alpha])
    L = np.zeros((3, 3, nwz), order="F")
    sig = np.zeros((3, nwz), order="F")
    for c in range(3):
        L[c, :, :] = scales[c]
        sig[c, :] = intensities[c] * u_ref
    # Davenport-type coherence, rho_ij**|f| = exp(-decay*|dz|*|f|/Ubar)
    dz = np.abs(z[:, None] - z[None, :])
    ubar = 0.5 * (U[:, None] + U[None, :])
    rho = np.exp(-decay * dz / ubar)
    packed = pack_corr(rho)
    nod_corr = np.asfortranarray(np.stack([packed] * 3, axis=1))
    # quasi-steady drag F = 1/2 rho Cd A (U+u)^2 -> linear rho Cd A U, quadratic 1/2 rho Cd A
    libs_load = np.asarray(libs_load, dtype=np.int32)
    wfc = np.zeros((libs_load.size, 6, nn), order="F")
    comp_gain = {1: 1.0, 2: 0.5, 3: 0.25}
    for li, lib in enumerate(libs_load):
        arm = 1.0 if lib == 1 else 0.4
        for c in comps:
            g = comp_gain[c] * arm
            wfc[li, c - 1, :] = g * RHO_AIR * cd * area * U
            wfc[li, c + 2, :] = g * 0.5 * RHO_AIR * cd * area
    if natf is None:
        natf = np.array([0.2, 0.6, 1.2])[:nm] if nm <= 3 else np.geomspace(0.15, 3.0, nm)
    natf = np.asarray(natf, dtype=np.float64)
    if xi is None:
        xi = np.full(nm, 0.01)
    xi = np.asarray(xi, dtype=np.float64)
    phi = np.empty((nn * nlibs, nm))
    for m in range(nm):
        phi[0::nlibs, m] = (z / height) ** (1.5 * (m + 1)) + rng.normal(0.0, 0.02, nn)
        for lib in range(1, nlibs):
            phi[lib::nlibs, m] = 0.3 * np.sin((m + lib) * np.pi * z / (2 * height)) + \
                rng.normal(0.0, 0.02, nn)
    phi = _normalise_modes(phi)
    Mm, Km, Cm = _modal_matrices(natf, xi)
    return Problem(nn=nn, nlibs=nlibs, n_load=np.arange(1, nn + 1, dtype=np.int32),
                   libs_load=libs_load, phi=phi, natf=natf, Mm=Mm, Km=Km,
                   Cm=Cm, psd_id=psd_id, tc=np.array(comps, dtype=np.int32), wz_node=wz_node,
                   u_node=U, turb_scales=L, turb_std=sig, u_mean_ref=u_ref,
                   nod_corr=nod_corr, wfc=wfc, nfreqs=nfreqs, df=df, name="tower", coords=z,
                   meta=dict(seed=seed, xi=xi))


def deck(nn: int = 1000, nm: int = 30, comps=(1, 3), span: float = 1000.0,
         height: float = 50.0, nfreqs: int = 513, df: float = 1.0 / 128.0, seed: int = 0,
         fmin: float = 0.1, fmax: float = 2.0, xi: float = 0.005, u10: float = 30.0,
         alpha: float = 0.16, psd_id: int = 1, decay: float = 10.0,
         intensities=(0.15, 0.10, 0.08), scales=(100.0, 40.0, 20.0), cd: float = 0.8,
         clp: float = 3.0, width: float = 30.0, noise: float = 0.02) -> Problem:
    """Long-span bridge deck (configuration 3 by default: NN=1000 nodes over
    1000 m at 50 m, u+w, Cd=0.8, Cl'=3.0, B=30 m, NM=30 sinusoidal modes,
    f_n 0.1-2 Hz, 0.5 % damping, 1025-point grid over [-4, 4] Hz)."""
    rng = np.random.default_rng(seed)
    x = np.arange(1, nn + 1, dtype=np.float64) * span / nn
    U0 = u10 * (height / 10.0) ** alpha
    U = np.full(nn, U0)
    L = np.zeros((3, 3, 1), order="F")
    sig = np.zeros((3, 1), order="F")
    for c in range(3):
        L[c, :, 0] = scales[c]
        sig[c, 0] = intensities[c] * U0
    dx = np.abs(x[:, None] - x[None, :])
    rho = np.exp(-decay * dx / U0)
    packed = pack_corr(rho)
    nod_corr = np.asfortranarray(np.stack([packed] * 3, axis=1))
    comps = tuple(int(c) for c in comps)
    # vertical force per unit length*tributary: lift slope on w, drag-like term on u
    trib = span / nn
    wfc = np.zeros((1, 6, nn), order="F")
    for c in comps:
        if c == 1:
            wfc[0, 0, :] = RHO_AIR * cd * width * U0 * trib
            wfc[0, 3, :] = 0.5 * RHO_AIR * cd * width * trib
        elif c == 3:
            wfc[0, 2, :] = 0.5 * RHO_AIR * (clp + cd) * width * U0 * trib
            wfc[0, 5, :] = 0.05 * 0.5 * RHO_AIR * clp * width * trib
        else:
            wfc[0, 1, :] = 0.25 * RHO_AIR * cd * width * U0 * trib
            wfc[0, 4, :] = 0.25 * 0.5 * RHO_AIR * cd * width * trib
    natf = np.linspace(fmin, fmax, nm)
    xis = np.full(nm, xi)
    phi = np.empty((nn, nm))
    for k in range(nm):
        phi[:, k] = np.sin((k + 1) * np.pi * x / span) + rng.normal(0.0, noise, nn)
    phi = _normalise_modes(phi)
    Mm, Km, Cm = _modal_matrices(natf, xis)
    return Problem(nn=nn, nlibs=1, n_load=np.arange(1, nn + 1, dtype=np.int32),
                   libs_load=np.array([1], dtype=np.int32), phi=phi, natf=natf, Mm=Mm, Km=Km,
                   Cm=Cm, psd_id=psd_id, tc=np.array(comps, dtype=np.int32),
                   wz_node=np.ones(nn, dtype=np.int32), u_node=U, turb_scales=L,
                   turb_std=sig, u_mean_ref=np.array([U0]), nod_corr=nod_corr, wfc=wfc,
                   nfreqs=nfreqs, df=df, name="deck", coords=x, meta=dict(seed=seed, xi=xis))


def config1(seed: int = 0) -> Problem:
    """Plain Classic BFM, CPU-runnable: tower NN=20, NM=3, u only, 129^2 grid."""
    return tower(nn=20, nm=3, comps=(1,), nfreqs=65, df=1.0 / 32.0, seed=seed)


def config2(seed: int = 0) -> Problem:
    """Tower NN=200, u/v/w, NM=10 (f_n log-spaced 0.15-3 Hz, damping 0.5-2 %), 513^2 grid."""
    rng = np.random.default_rng(seed + 1000)
    xi = rng.uniform(0.005, 0.02, 10)
    return tower(nn=200, nm=10, comps=(1, 2, 3), natf=np.geomspace(0.15, 3.0, 10), xi=xi,
                 nfreqs=257, df=1.0 / 64.0, seed=seed)


def config3(seed: int = 0) -> Problem:
    """Bridge deck NN=1000, NM=30, u+w, 1025^2 grid over [-4, 4] Hz."""
    return deck(seed=seed)


def config5(seed: int = 0) -> Problem:
    """Bridge deck NN=500, NM=20 (the adaptive-mesher case)."""
    return deck(nn=500, nm=20, nfreqs=257, df=1.0 / 64.0, seed=seed)


def batch(n: int = 256, seed: int = 0, max_nn: int = 800):
    """Uneven structures: NN 50..800, NM 4..24, 1-3 components, grid side 257..1025,
    tower or deck geometry, U_ref 20..45 m/s."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n):
        nn = int(rng.integers(50, max_nn + 1))
        nm = int(rng.integers(4, 25))
        ntc = int(rng.integers(1, 4))
        comps = tuple(sorted(rng.choice([1, 2, 3], size=ntc, replace=False).tolist()))
        nfh = int(rng.choice([129, 257, 513]))
        uref = float(rng.uniform(20.0, 45.0))
        if rng.random() < 0.5:
            p = tower(nn=nn, nm=nm, comps=comps, natf=np.geomspace(0.15, 3.0, nm),
                      xi=rng.uniform(0.005, 0.02, nm), nfreqs=nfh, df=4.0 / (nfh - 1),
                      seed=seed * 1000 + s, u10=uref / (100.0 / 10.0) ** 0.16)
        else:
            p = deck(nn=nn, nm=nm, comps=comps, nfreqs=nfh, df=4.0 / (nfh - 1),
                     seed=seed * 1000 + s, u10=uref / (50.0 / 10.0) ** 0.16)
        out.append(p)
    return out
