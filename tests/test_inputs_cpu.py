"""CPU tests of the standalone program's input files (bsalib_amd.inputs):
writer / reader round trip, and the Fortran front's bsa_generateBSAInputFiles
against the Python writer for the same problem (two independent writers of
the layout the reference program reads, app/bsa.F90:642-930)."""
import numpy as np
import pytest

from bsalib_amd import inputs
from tests import fdriver
from tests.cases import CASES


@pytest.mark.parametrize("name", ["tiny_uvw", "config1"])
def test_round_trip(tmp_path, name):
    p = CASES[name]()
    inputs.write_inputs(p, tmp_path, settings=dict(analysis=3, isvd=0, bkgrfmt=33))
    b = inputs.read_bsadata(tmp_path / inputs.BSA_DATA_FNAME)
    assert (b["analysis"], b["isvd"], b["bkgrfmt"], b["nfreqs"]) == (3, 0, 33, p.nfreqs)
    assert b["df"] == p.df and b["tc"] == list(p.tc)
    e = inputs.read_extdata(tmp_path / inputs.EXT_DATA_FNAME)
    assert e["nn"] == p.nn and e["nm"] == p.phi.shape[1] and e["nz"] == p.nwz
    for k in ("n_load", "libs_load", "wz_node"):
        assert np.array_equal(e[k], getattr(p, k))
    for k in ("u_node", "u_mean_ref", "turb_scales", "turb_std", "wfc", "phi", "natf", "Mm",
              "Km", "Cm"):
        assert np.array_equal(e[k], getattr(p, k)), k
    assert np.array_equal(e["nod_corr"], p.nod_corr[:, :3])


@pytest.mark.parametrize("name", ["tiny_uvw", "tiny_dav"])
def test_front_generated_files_match_python_writer(tmp_path, name):
    p = CASES[name]()
    out, r = fdriver.run(p, "gen", workdir=tmp_path)   # bsa_Run returns after writing
    assert "Generating BSA-compatible input files" in r.stdout
    (tmp_path / "py").mkdir()
    inputs.write_inputs(p, tmp_path / "py", settings=dict(analysis=1, isvd=1))
    bf = inputs.read_bsadata(tmp_path / inputs.BSA_DATA_FNAME)
    bp = inputs.read_bsadata(tmp_path / "py" / inputs.BSA_DATA_FNAME)
    assert bf == bp
    ef = inputs.read_extdata(tmp_path / inputs.EXT_DATA_FNAME)
    ep = inputs.read_extdata(tmp_path / "py" / inputs.EXT_DATA_FNAME)
    assert ef.keys() == ep.keys()
    for k in ef:
        if k == "xsi":      # the driver derives it from C, K and M
            assert np.allclose(ef[k], ep[k], rtol=1e-14)
        else:
            assert np.array_equal(np.asarray(ef[k]), np.asarray(ep[k])), k


def test_bisp_export_reader_round_trip(tmp_path):
    from bsalib_amd import exports
    nf = 5
    f = (np.arange(nf) - 2) * 0.25
    vals = np.arange(nf * nf, dtype=np.float64).reshape(nf, nf, order="F") * 1e-3
    head = np.array([3, 2, 1, 3, 1, 0, 1, nf * nf, 1, nf, nf], dtype=np.int32)
    rec = np.stack([np.tile(f, nf), np.repeat(f, nf), vals.reshape(-1, order="F")], 1)
    (tmp_path / "x.brm").write_bytes(head.tobytes() + rec.astype(np.float32).tobytes())
    e = exports.read_bisp_export(tmp_path / "x.brm")
    assert e["nm"] == 3 and list(e["indexes"]) == [2, 1, 3]
    assert (e["ncomb"], e["ispsym"], e["nzones"], e["npts"]) == (1, 0, 1, nf * nf)
    (z,) = e["zones"]
    assert (z["id"], z["ni"], z["nj"]) == (1, nf, nf)
    assert np.array_equal(z["values"][:, 0], vals.reshape(-1, order="F").astype(np.float32))
    (tmp_path / "y.brm").write_bytes(head.tobytes() + rec.astype(np.float32).tobytes() + b"\0")
    with pytest.raises(ValueError):
        exports.read_bisp_export(tmp_path / "y.brm")


def test_reference_module_extras_without_device(tmp_path):
    """Procedures and constants of the reference's module BsaLib that need no
    device (src/BsaLib.F90: bsa_exportPSDToFile, bsa_exportBispToFile,
    bsa_isFullComp, bsa_setPremeshMode, bsa_setPolicyIDValidationValues,
    bsa_setOutUnit, the CST_* and export-name constants): a program using them
    compiles against module BsaLib and writes the reference's text layouts."""
    import subprocess
    from bsalib_amd import build
    fdriver.ensure_built()
    src = build.ROOT / "tests" / "fortran" / "api_extras.F90"
    exe = tmp_path / "api_extras"
    subprocess.run([build.FC, "-cpp", "-O2", "-I", str(build.FOUT), str(src), str(build.FLIB),
                    "-L", str(build.OUT_DIR), "-lbsahip", f"-Wl,-rpath,{build.OUT_DIR}",
                    "-o", str(exe)], check=True, cwd=tmp_path, capture_output=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert [x.split()[1] for x in lines if x.startswith("FULL")] == ["T", "F"]
    cst = [x for x in lines if x.startswith("CST")][0].split()
    assert float(cst[1]) == 2 * np.pi and abs(float(cst[2]) - 1.5 * np.pi) < 1e-15
    assert [x for x in lines if x.startswith("NAMES")][0].split()[1:] == ["m3mr_cls", "bsaexport_msh.brm"]
    assert "pre-meshing mode value" in r.stdout
    psd = np.arange(1, 7).reshape(3, 2, order="F") * 0.5
    f = np.array([0.0, 0.1, 0.2])
    v = (tmp_path / "psd.txt").read_text().split()
    assert (int(v[0]), int(v[1])) == (3, 2)
    assert np.allclose(np.array(v[2:], float).reshape(3, 3), np.column_stack([f, psd]), rtol=1e-15)
    v = (tmp_path / "psd_nof.txt").read_text().split()
    assert np.allclose(np.array(v[2:], float).reshape(3, 2), psd, rtol=1e-15)
    bisp = np.arange(1, 13).reshape(2, 2, 3, order="F") * 0.25
    v = (tmp_path / "bisp.txt").read_text().split()
    assert [int(x) for x in v[:3]] == [2, 2, 3]
    rows = np.array(v[3:], float).reshape(6, 2)
    want = np.array([bisp[:, i, j] for j in range(3) for i in range(2)])
    assert np.allclose(rows, want, rtol=1e-15)
