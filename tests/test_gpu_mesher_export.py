"""Bispectrum export and checkpoint of the device Mesher.

* bsahip_mesher_export: at every tile's post-mesh points, the values equal the
  CPU restatement's integrand (getFM_full_tnm_scalar_msh_ BFM, bilinear
  interpolation, FRF product) at 1e-12, the point frequencies are the
  restatement's bit for bit, and the trapezoid-weighted sum of the exported
  values is each tile's m3mr contribution of bsahip_mesher_run_tiles (1e-12).
  Unit coefficient columns select entries exactly; other columns are the
  linear combinations of the entries.
* bsahip_mesher_save / bsahip_mesher_load: a Mesher rebuilt from its
  checkpoint exports bit for bit like the original and runs to the same
  moments (up to the order of the atomic sums); a checkpoint of
  another layout, a truncated file or a foreign file is refused.
* The Fortran front in visual mode (modal and nodal indexes) and with
  bsa_setBRMExportDefaultMode(BASE) writes the reference's stream layout with
  one zone per tile; its values are the C ABI export rounded to real32.

The cases use 10-20 % modal damping so that the post-mesh grids stay small.
"""
import copy

import numpy as np
import pytest

from bsalib_amd import abi, exports, host
from oracle import oracle as O
from tests import fdriver
from tests.cases import CASES
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _damped(name, scale):
    p = copy.deepcopy(CASES[name]())
    p.Cm = p.Cm * scale
    return p


def _tile_weights(t, dims):
    na, nb = int(dims[0]) - 1, int(dims[1]) - 1
    h1, h2 = (t[1] - t[0]) / na, (t[3] - t[2]) / nb
    w1 = np.ones(na + 1); w1[[0, -1]] = 0.5
    w2 = np.ones(nb + 1); w2[[0, -1]] = 0.5
    return (np.outer(w1, w2) * (2 * np.pi * h1) * (2 * np.pi * h2)).reshape(-1, order="F")


def test_export_equals_oracle_and_sums_to_tile_moments(gpu_lib):
    p = _damped("tiny_u", 10)
    M = abi.Mesher(p, bkg_base_rfmnt=5, tile_cells=2, bkg_peakw=0.25)
    geom, _ = M.tiles()
    lv = np.random.default_rng(5).choice([1, 2, 4], size=M.ntiles).astype(np.int32)
    e = M.export(levels=lv)
    r = M.run(levels=lv, per_tile=True)
    M.close()
    tiles = []
    O.mesher_integrate(p, geom, lv, brm_out=tiles)
    nm3 = M.nm ** 3
    assert e["val"].shape == (e["fi"].size, nm3)
    off = 0
    for k, (t, d, o) in enumerate(zip(geom, e["dims"], tiles)):
        n = int(d[0]) * int(d[1])
        assert (int(d[0]), int(d[1])) == (o["f1"].size, o["f2"].size)
        assert np.array_equal(e["fi"][off:off + n], np.tile(o["f1"], o["f2"].size))
        assert np.array_equal(e["fj"][off:off + n], np.repeat(o["f2"], o["f1"].size))
        v = e["val"][off:off + n]
        assert rel_err(v, o["brm"].reshape(nm3, -1, order="F").T) < 1e-12, k
        assert rel_err(_tile_weights(t, d) @ v, r["tile_m3mr"][k]) < 1e-12, k
        off += n
    assert off == e["fi"].size == r["brm_points"]


def test_export_coefficient_columns(gpu_lib):
    p = _damped("tiny_uvw", 20)
    M = abi.Mesher(p, bkg_base_rfmnt=5, tile_cells=2, bkg_peakw=0.25)
    nm3 = M.nm ** 3
    full = M.export()["val"]
    rng = np.random.default_rng(8)
    sel = [0, nm3 - 1, 3]
    coef = np.zeros((nm3, 5))
    for c, q in enumerate(sel):
        coef[q, c] = 1.0
    phi = p.phi[:, M.modes - 1]
    coef[:, 3] = np.einsum("m,n,p->mnp", phi[2], phi[2], phi[2]).reshape(-1, order="F")
    coef[:, 4] = rng.normal(size=nm3)
    got = M.export(coef=coef)["val"]
    for c, q in enumerate(sel):
        assert np.array_equal(got[:, c], full[:, q])
    assert rel_err(got[:, 3:], full @ coef[:, 3:]) < 1e-12
    # size query, capacity and argument checks
    lib = abi.lib()
    import ctypes as C
    n = C.c_int64(0)
    abi._check(lib.bsahip_mesher_export(M.h, None, 0, None, 0, None, None, None, None, C.byref(n)))
    assert n.value == full.shape[0]
    small = np.zeros(4)
    with pytest.raises(abi.BsaHipError) as ei:
        abi._check(lib.bsahip_mesher_export(M.h, None, nm3, None, 1, None, None, None,
                                            abi._ptr(small, C.c_double), C.byref(n)))
    assert ei.value.code == abi.EINVAL
    with pytest.raises(abi.BsaHipError):
        M.export(levels=3)
    M.close()


def test_checkpoint_round_trip(gpu_lib, tmp_path):
    p = _damped("tiny_u", 10)
    kw = dict(bkg_base_rfmnt=5, tile_cells=2, bkg_peakw=0.25)
    M = abi.Mesher(p, **kw)
    lv = np.random.default_rng(2).choice([1, 2, 4, 8], size=M.ntiles).astype(np.int32)
    path = tmp_path / "zones.dump"
    M.save(path)
    a, ea = M.run(levels=lv), M.export(levels=lv, coef=np.eye(M.nm ** 3)[:, :2])
    M.close()
    L = abi.Mesher(p, load=path, **kw)
    b, eb = L.run(levels=lv), L.export(levels=lv, coef=np.eye(L.nm ** 3)[:, :2])
    L.close()
    # the moments are summed with atomics (order varies run to run); the
    # exported values are written point by point
    for k in ("m3mf", "m3mr"):
        assert rel_err(a[k], b[k]) < 1e-13
    assert (a["evals_pre"], a["evals_post"], a["brm_points"]) == \
        (b["evals_pre"], b["evals_post"], b["brm_points"])
    assert np.array_equal(ea["val"], eb["val"])
    # another layout (other tile count; same count, other geometry), a
    # truncated file, a foreign file
    for other in (dict(kw, bkg_base_rfmnt=7), dict(kw, bkg_peakw=0.3)):
        with pytest.raises(abi.BsaHipError) as ei:
            abi.Mesher(p, load=path, **other)
        assert ei.value.code == abi.EINVAL
    raw = path.read_bytes()
    (tmp_path / "cut.dump").write_bytes(raw[:-8])
    with pytest.raises(abi.BsaHipError):
        abi.Mesher(p, load=tmp_path / "cut.dump", **kw)
    (tmp_path / "junk.dump").write_bytes(b"\0" * 256)
    with pytest.raises(abi.BsaHipError):
        abi.Mesher(p, load=tmp_path / "junk.dump", **kw)


def _front_export(p, tmp_path, *opts, nval=1, fname="bsaexport_msh.brm"):
    out, r = fdriver.run(p, "mesher", *opts, workdir=tmp_path)
    assert "Exporting  bispectrum" in r.stdout
    return out, exports.read_bisp_export(tmp_path / fname, nval=nval)


def _front_mesher(p):
    return abi.Mesher(p, bkg_base_rfmnt=20, tile_cells=32, bkg_peakw=host.bkg_peak_width(p))


def _same_values(zones, e, ncol):
    off = 0
    tol_abs = 1e-12 * np.abs(e["val"]).max()
    for z, d in zip(zones, e["dims"]):
        n = int(d[0]) * int(d[1])
        assert (z["id"], z["ni"], z["nj"]) == (1, int(d[0]), int(d[1]))
        assert np.array_equal(z["fi"], e["fi"][off:off + n].astype(np.float32))
        assert np.array_equal(z["fj"], e["fj"][off:off + n].astype(np.float32))
        w32 = e["val"][off:off + n, :ncol].astype(np.float32)
        tol = np.spacing(np.abs(w32)).astype(np.float64) + tol_abs
        assert np.all(np.abs(z["values"].astype(np.float64) - w32.astype(np.float64)) <= tol)
        off += n
    assert off == e["fi"].size


def test_front_mesher_visual_modal_and_nodal(gpu_lib, tmp_path):
    p = _damped("tiny_u", 20)
    M = _front_mesher(p)
    lv, _ = M.adapt(1e-3)                 # the front's default refinement
    nm = M.nm
    idx = (1, nm, 2)
    q = idx[0] - 1 + (idx[1] - 1) * nm + (idx[2] - 1) * nm * nm
    unit = np.zeros((nm ** 3, 1))
    unit[q, 0] = 1.0
    em = M.export(levels=lv, coef=unit)
    d1 = tmp_path / "modal"
    d1.mkdir()
    out, f = _front_export(p, d1, "visual=%d,%d,%d" % idx)
    ref = M.run(levels=lv)
    assert rel_err(out["m3mr_msh"], ref["m3mr"]) < 1e-13          # moments unchanged
    assert f["nm"] == 3 and list(f["indexes"]) == list(idx)
    assert (f["ncomb"], f["ispsym"], f["nzones"], f["npts"]) == (1, 0, M.ntiles, em["fi"].size)
    _same_values(f["zones"], em, 1)
    # nodal: node 4, DOF 1
    d = (4 - 1) * p.nlibs + 1
    ph = p.phi[d - 1, M.modes - 1]
    coef = np.einsum("m,n,p->mnp", ph, ph, ph).reshape(-1, 1, order="F")
    en = M.export(levels=lv, coef=coef)
    d2 = tmp_path / "nodal"
    d2.mkdir()
    _, f = _front_export(p, d2, "visualn=4,1", "sym=4", fname="bsaexport_msh.brn")
    assert f["nm"] == 2 and list(f["indexes"]) == [4, 1]
    assert (f["ncomb"], f["ispsym"], f["nzones"]) == (1, 4, M.ntiles)
    _same_values(f["zones"], en, 1)
    M.close()


def test_front_mesher_brm_export_base(gpu_lib, tmp_path):
    p = _damped("tiny_uvw", 20)
    M = _front_mesher(p)
    nm3 = M.nm ** 3
    e = M.export(levels=M.adapt(1e-3)[0])
    M.close()
    _, f = _front_export(p, tmp_path, "brmbase", nval=nm3)
    assert f["nm"] == M.nm and list(f["indexes"]) == list(M.modes)
    assert (f["ncomb"], f["nzones"], f["npts"]) == (nm3, M.ntiles, e["fi"].size)
    _same_values(f["zones"], e, nm3)


def _callback(stdout):
    line = [x for x in stdout.splitlines() if x.startswith("CALLBACK")]
    assert line, stdout[-2000:]
    _, zones, pts, total = line[0].split()
    return int(zones), int(pts), float(total)


def test_front_export_callback_replaces_files(gpu_lib, tmp_path):
    """bsa_setBispExportCallback: the user routine receives, per zone, the
    values the export file would hold (and no file is written)."""
    from tests.conftest import load_golden
    p = CASES["tiny_u"]()
    g = load_golden("tiny_u")
    nm = 3
    d1 = tmp_path / "cls"
    d1.mkdir()
    _, r = fdriver.run(p, "visual=3,1,2", "cb", workdir=d1)
    zones, pts, total = _callback(r.stdout)
    want = g["bisp_r"][:, :, 3 - 1 + (1 - 1) * nm + (2 - 1) * nm * nm]
    assert (zones, pts) == (1, p.nf ** 2)
    assert abs(total - want.sum()) <= 1e-12 * np.abs(want).sum()
    assert not (d1 / "bsaexport_cls.brm").exists()
    q = _damped("tiny_uvw", 20)
    M = _front_mesher(q)
    e = M.export(levels=M.adapt(1e-3)[0])
    M.close()
    d2 = tmp_path / "msh"
    d2.mkdir()
    _, r = fdriver.run(q, "mesher", "brmbase", "cb", workdir=d2)
    zones, pts, total = _callback(r.stdout)
    assert (zones, pts) == (M.ntiles, e["fi"].size)
    assert abs(total - e["val"].sum()) <= 1e-12 * np.abs(e["val"]).sum()
    assert not (d2 / "bsaexport_msh.brm").exists()


def test_front_only_premesh_and_modal_data(gpu_lib, tmp_path):
    """bsa_enableOnlyPremesh stops the Mesher after its pre-mesh (no Mesher
    outputs); bsa_exportModalData writes modal.txt at bsa_Run."""
    p = _damped("tiny_uvw", 20)
    out, r = fdriver.run(p, "mesher", "premesh", "modal", workdir=tmp_path)
    assert "Skipping Post-meshing phase!" in r.stdout
    assert out["m3mf_msh"] is None and out["m3mr_msh"] is None
    v = (tmp_path / "modal.txt").read_text().split()
    nn, nlibs, nm = int(v[0]), int(v[1]), int(v[2])
    assert (nn, nlibs, nm) == (p.nn, p.nlibs, p.phi.shape[1])
    vals = np.array(v[3:], dtype=float).reshape(nm, 1 + nn * nlibs)
    assert np.allclose(vals[:, 0], p.natf, rtol=1e-9)
    assert np.allclose(vals[:, 1:].T, p.phi, rtol=1e-9, atol=1e-12)
