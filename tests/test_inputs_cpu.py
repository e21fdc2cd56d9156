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
