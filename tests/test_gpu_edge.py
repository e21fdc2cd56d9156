"""GPU parity at the smallest shapes the Classic path accepts: one loaded node,
one mode, the 3-point grid (nfreqs = 2), and a single mode over many nodes
(K padded far beyond the data).  FP64 against the oracle at 1e-12 normwise;
FP32 and split-bf16x3 at 1e-4.  Diagonal model and point bispectra on the
same shapes."""
import numpy as np
import pytest

from bsalib_amd import synth
from oracle import oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

SHAPES = {
    "one_node_one_mode_3pt": dict(nn=1, nm=1, nfreqs=2),
    "three_nodes_uvw_one_mode": dict(nn=3, nm=1, nfreqs=3, comps=(1, 2, 3)),
    "five_nodes_two_modes_3pt": dict(nn=5, nm=2, nfreqs=2),
    "many_nodes_one_mode": dict(nn=200, nm=1, nfreqs=9),
    "two_nodes_three_modes": dict(nn=2, nm=3, nfreqs=5, comps=(1, 3)),
}


def _prob(name):
    return synth.tower(seed=11, df=0.125, **SHAPES[name])


@pytest.mark.parametrize("name", list(SHAPES))
def test_small_shapes_fp64(gpu_lib, name):
    p = _prob(name)
    r = gpu_lib.classic_run(p)
    o = O.classic_factorised(p)
    assert rel_err(r["m3mf"], o["m3mf"]) < 1e-12
    assert rel_err(r["m3mr"], o["m3mr"]) < 1e-12
    m = O.modal_psd(p)
    for k in ("m2mf", "m2mr", "m2o2mr"):
        assert rel_err(r[k], m[k]) < 1e-12, k


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("name", list(SHAPES))
def test_small_shapes_mixed(gpu_lib, name, prec):
    p = _prob(name)
    r = gpu_lib.classic_run(p, prec=prec)
    o = O.classic_factorised(p)
    assert rel_err(r["m3mf"], o["m3mf"]) < 1e-4
    assert rel_err(r["m3mr"], o["m3mr"]) < 1e-4


@pytest.mark.parametrize("name", list(SHAPES))
def test_small_shapes_spectra_and_points(gpu_lib, name):
    p = _prob(name)
    s = gpu_lib.classic_spectra(p)
    o = O.classic_factorised(p, want_bisp=True)
    nf = p.nf
    B = s["bisp_f"].reshape(nf * nf, -1, order="F").T
    assert rel_err(B, o["bfm"]) < 1e-12
    f = p.freqs()
    fi = np.repeat(f, nf)
    fj = np.tile(f, nf)
    bfm, _ = gpu_lib.bfm_points_cls(p, fi, fj, with_brm=False)
    want = s["bisp_f"].reshape(nf * nf, -1, order="C").T
    assert rel_err(bfm, want) < 1e-12


@pytest.mark.parametrize("name", list(SHAPES))
def test_small_shapes_diag(gpu_lib, name):
    p = _prob(name)
    r = gpu_lib.classic_run(p, only_diag=True)
    o = O.classic_diag(p)
    for k in ("m2mf", "m2mr", "m3mf", "m3mr"):
        assert rel_err(r[k], o[k]) < 1e-12, k
