"""The standalone program's input files, bsa.bsadata and bsa.extdata.

Layouts follow the reference program's reader (app/bsa.F90:642-930) and the
library's writer (src/bsa/BsaLibImpl.F90:2065-2155):

* bsa.bsadata -- text, five labelled sections, one value per line:
  -GENERAL: analysis type, version, PSD convention (1 pulsation, 0 frequency),
  PSD switch, bispectra switch, only-diagonal switch, spatial symmetry, tensor
  symmetry, test mode; -CLASSIC: Classic version (0 vectorised, 1 scalar),
  number of positive frequencies, delta f; -MESHER: POD switch, background base
  refinement, background / peak / max area extensions, full coverage, dump modal
  info; -DIRECTIONS: count then one direction per line; -TURBULENCE: count then
  one component per line.
* bsa.extdata -- a stream of int32 and real64 values (Fortran order):
  NN, NLIBS, NNL, NLIBSL, loaded nodes, loaded DOFs, coordinates (3, NN);
  vertical profile, PSD type, vertical axis, NDEGW, air density, rotation (3, 3),
  number of zones NZ, reference altitudes (NZ), mean speeds (NZ), L (3, 3, NZ),
  sigma (3, NZ), coherence coefficients (3, 3, NZ) and exponents (3, 3, NZ),
  zone limits (NZ+1), zone rotations (3, 3, NZ), incidence angles (NZ), nodal
  zones (NN), nodal altitudes (NN), nodal speeds (NN), nodal coherence
  ((NN^2+NN)/2, 3), wind force coefficients (NLIBSL, NDEGW+3, NNL); NDOFS, NM,
  natural frequencies (NM), mode shapes (NDOFS, NM), modal masses (NM),
  stiffnesses (NM), damping matrix (NM, NM), damping ratios (NM).

write_inputs writes both files for a problem (the generators' Problem type);
read_bsadata / read_extdata read them back.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

BSA_DATA_FNAME = "bsa.bsadata"
EXT_DATA_FNAME = "bsa.extdata"

_GENERAL = ("analysis", "version", "scaling", "psd", "bisp", "only_diag", "spatial_sym",
            "tensor_sym", "test")
_MESHER = ("isvd", "bkgrfmt", "bkgaext", "genpaext", "maxaext", "fcov", "dumpmod")

DEFAULT_SETTINGS = dict(analysis=1, version=2, scaling=1, psd=1, bisp=1, only_diag=0,
                        spatial_sym=0, tensor_sym=0, test=1, classic_mode=0, isvd=1,
                        bkgrfmt=20, bkgaext=2.0, genpaext=2.0, maxaext=2.0, fcov=1,
                        dumpmod=0, dirs=(1,))


def write_bsadata(path, settings: dict, nfreqs: int, df: float, tc):
    s = dict(DEFAULT_SETTINGS, **settings)
    lines = ["-GENERAL:"] + [f"{int(s[k]):8d}" for k in _GENERAL]
    lines += ["-CLASSIC:", f"{int(s['classic_mode']):8d}", f"{int(nfreqs):8d}", f"{float(df):25.17e}"]
    lines += ["-MESHER:"]
    for k in _MESHER:
        v = s[k]
        lines.append(f"{float(v):25.17e}" if k.endswith("aext") else f"{int(v):8d}")
    dirs = list(s["dirs"])
    lines += ["-DIRECTIONS:", f"{len(dirs):8d}"] + [f"{int(d):8d}" for d in dirs]
    lines += ["-TURBULENCE:", f"{len(tc):8d}"] + [f"{int(t):8d}" for t in tc]
    Path(path).write_text("\n".join(lines) + "\n")


def read_bsadata(path) -> dict:
    """Values of bsa.bsadata (section labels are checked, not interpreted)."""
    rows = [r.strip() for r in Path(path).read_text().splitlines()]
    it = iter(rows)

    def label(name):
        lab = next(it)
        if not lab.upper().startswith(name):
            raise ValueError(f"expected section {name}, found {lab!r}")

    out = {}
    label("-GENERAL")
    for k in _GENERAL:
        out[k] = int(next(it))
    label("-CLASSIC")
    out["classic_mode"] = int(next(it))
    out["nfreqs"] = int(next(it))
    out["df"] = float(next(it).replace("d", "e").replace("D", "e"))
    label("-MESHER")
    for k in _MESHER:
        v = next(it)
        out[k] = float(v.replace("d", "e").replace("D", "e")) if k.endswith("aext") else int(v)
    label("-DIRECTIONS")
    out["dirs"] = [int(next(it)) for _ in range(int(next(it)))]
    label("-TURBULENCE")
    out["tc"] = [int(next(it)) for _ in range(int(next(it)))]
    return out


def write_extdata(path, prob, coords=None):
    """bsa.extdata for a problem.  Data the device path does not read (nodal
    coordinates unless given, reference altitudes, coherence coefficients and
    exponents, zone limits, zone rotations, incidence angles, nodal altitudes)
    is written as zeros; air density 1.25 and an identity rotation."""
    nn, nz = int(prob.nn), int(prob.nwz)
    nnl, nlibsl = int(prob.nnl), int(prob.libs_load.size)
    i32 = lambda a: np.asarray(a, dtype=np.int32).ravel(order="F").tobytes()
    f64 = lambda a: np.asarray(a, dtype=np.float64).ravel(order="F").tobytes()
    corr = np.zeros((prob.nod_corr.shape[0], 3))
    corr[:, :min(3, prob.nod_corr.shape[1])] = prob.nod_corr[:, :3]
    xi = np.asarray(prob.meta["xi"], dtype=np.float64)
    parts = [
        i32([nn, prob.nlibs, nnl, nlibsl]), i32(prob.n_load), i32(prob.libs_load),
        f64(np.zeros((3, nn)) if coords is None else np.asarray(coords).reshape(3, nn, order="F")),
        i32([1, prob.psd_id, 3, prob.wfc.shape[1] - 3]), f64([1.25]), f64(np.eye(3)), i32([nz]),
        f64(np.zeros(nz)), f64(prob.u_mean_ref), f64(prob.turb_scales), f64(prob.turb_std),
        f64(np.zeros((3, 3, nz))), f64(np.zeros((3, 3, nz))), f64(np.zeros(nz + 1)),
        f64(np.zeros((3, 3, nz))), f64(np.zeros(nz)), i32(prob.wz_node), f64(np.zeros(nn)),
        f64(prob.u_node), f64(corr), f64(prob.wfc),
        i32([prob.phi.shape[0], prob.phi.shape[1]]), f64(prob.natf), f64(prob.phi), f64(prob.Mm),
        f64(prob.Km), f64(prob.Cm), f64(xi),
    ]
    Path(path).write_bytes(b"".join(parts))


def read_extdata(path) -> dict:
    """Arrays of bsa.extdata, by the reference reader's names."""
    buf = memoryview(Path(path).read_bytes())
    pos = 0

    def take(dtype, shape):
        nonlocal pos
        n = int(np.prod(shape))
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=pos)
        pos += n * np.dtype(dtype).itemsize
        return a.reshape(shape, order="F").copy() if len(shape) > 1 else a.copy()

    d = {}
    d["nn"], d["nlibs"], d["nnl"], d["nlibsl"] = (int(v) for v in take(np.int32, (4,)))
    nn, nnl, nlibsl = d["nn"], d["nnl"], d["nlibsl"]
    d["n_load"] = take(np.int32, (nnl,))
    d["libs_load"] = take(np.int32, (nlibsl,))
    d["coords"] = take(np.float64, (3, nn))
    d["wind_prof"], d["psd_id"], d["vert"], d["ndegw"] = (int(v) for v in take(np.int32, (4,)))
    d["air_density"] = float(take(np.float64, (1,))[0])
    d["rot_w2g"] = take(np.float64, (3, 3))
    nz = d["nz"] = int(take(np.int32, (1,))[0])
    d["z_ref"] = take(np.float64, (nz,))
    d["u_mean_ref"] = take(np.float64, (nz,))
    d["turb_scales"] = take(np.float64, (3, 3, nz))
    d["turb_std"] = take(np.float64, (3, nz))
    d["corr_coeffs"] = take(np.float64, (3, 3, nz))
    d["corr_expnts"] = take(np.float64, (3, 3, nz))
    d["zone_limits"] = take(np.float64, (nz + 1,))
    d["rot_zones"] = take(np.float64, (3, 3, nz))
    d["incidence"] = take(np.float64, (nz,))
    d["wz_node"] = take(np.int32, (nn,))
    d["altitude"] = take(np.float64, (nn,))
    d["u_node"] = take(np.float64, (nn,))
    d["nod_corr"] = take(np.float64, ((nn * nn + nn) // 2, 3))
    d["wfc"] = take(np.float64, (nlibsl, d["ndegw"] + 3, nnl))
    d["ndofs"], d["nm"] = (int(v) for v in take(np.int32, (2,)))
    nm = d["nm"]
    d["natf"] = take(np.float64, (nm,))
    d["phi"] = take(np.float64, (d["ndofs"], nm))
    d["Mm"] = take(np.float64, (nm,))
    d["Km"] = take(np.float64, (nm,))
    d["Cm"] = take(np.float64, (nm, nm))
    d["xsi"] = take(np.float64, (nm,))
    if pos != len(buf):
        raise ValueError(f"{path}: {len(buf) - pos} trailing bytes")
    return d


def write_inputs(prob, directory, settings: dict | None = None, coords=None):
    """Both input files for a problem into directory (its Classic grid, its
    turbulence components, the given settings over DEFAULT_SETTINGS)."""
    directory = Path(directory)
    write_bsadata(directory / BSA_DATA_FNAME, settings or {}, prob.nfreqs, prob.df, prob.tc)
    write_extdata(directory / EXT_DATA_FNAME, prob, coords)
