"""GPU parity of the zone Mesher (bsahip_mesher_*):
* one background zone sampled at the BRM spacing is the tensor trapezoid rule on a
  uniform grid, so it must equal the Classic G path with trapezoid weights and
  S(f1+f2) evaluated exactly (a different kernel) -- 1e-12;
* on a multi-tile layout with mixed refinement levels it must equal the CPU
  restatement (getFM_full_tnm_scalar_msh_ per BFM point, bilinear interpolation,
  FRF, trapezoid weights) -- 1e-12, with the same BFM evaluation counts;
* on 256 tiles of five policies whose BRM grids share one lattice, with the BFM
  at every BRM point, the tiles' trapezoid rules sum to the tensor trapezoid of
  the union grid: it must equal the Classic G path (trapezoid, S(f1+f2) exact)
  on that grid, and so must the CPU restatement -- 1e-12 (seams and edge weights
  pinned by mathematics, not by a restatement of the same design);
* fully refined, it converges to the Classic integral on a fine uniform grid."""
import numpy as np
import pytest

from bsalib_amd import abi
from oracle import oracle as O
from tests.cases import CASES, LATTICE_CLASSIC, LATTICE_MESHER
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def test_single_zone_equals_classic_trapezoid(gpu_lib):
    p = CASES["tiny_uvw"]()                      # f_n = 0.2, 0.6 Hz
    # b = 1.5 >= fmax = 1.2: one background tile of 8 x 8 level-0 cells over [-1.5, 1.5]^2;
    # wide layout peaks (w = 0.8 and 1.2, inside b) keep 8 BRM cells per level-0 cell, so
    # level 8 puts the BFM on the BRM grid: h = 3/64
    M = abi.Mesher(p, bkg_base_rfmnt=9, tile_cells=8, bkg_peakw=1.5, xsi=[1.0, 0.5])
    assert M.ntiles == 1
    r = M.run(levels=8)
    M.close()
    q = CASES["tiny_uvw"]()
    q.nfreqs, q.df = 33, 3.0 / 64.0
    c = abi.classic_run(q, quad="trapezoid", exact=True)
    assert rel_err(r["m3mf"], c["m3mf"]) < 1e-12
    assert rel_err(r["m3mr"], c["m3mr"]) < 1e-12
    assert r["evals_pre"] == 81 and r["evals_post"] == 65 * 65 - 81
    assert r["brm_points"] == 65 * 65


def test_lattice_layout_equals_classic_trapezoid_and_oracle(gpu_lib):
    p = CASES["tiny_uvw"]()
    M = abi.Mesher(p, **LATTICE_MESHER)
    geom, _ = M.tiles()
    assert M.ntiles == 256 and set(geom[:, 6].astype(int)) == {3, 4, 5, 9, 100}
    r = M.run(levels=2)
    M.close()
    q = CASES["tiny_uvw"]()
    q.nfreqs, q.df = LATTICE_CLASSIC["nfreqs"], LATTICE_CLASSIC["df"]
    c = abi.classic_run(q, quad="trapezoid", exact=True)
    assert rel_err(r["m3mf"], c["m3mf"]) < 1e-12
    assert rel_err(r["m3mr"], c["m3mr"]) < 1e-12
    o = O.mesher_integrate(p, geom, 2)
    assert rel_err(r["m3mf"], o["m3mf"]) < 1e-12
    assert rel_err(r["m3mr"], o["m3mr"]) < 1e-12
    assert (r["evals_pre"], r["evals_post"]) == (o["evals_pre"], o["evals_post"])
    assert r["evals_pre"] == 43 ** 2


def test_multi_tile_mixed_levels_match_oracle(gpu_lib):
    p = CASES["tiny_uvw"]()
    M = abi.Mesher(p, bkg_base_rfmnt=5, tile_cells=2, bkg_peakw=0.25)
    geom, rule = M.tiles()
    assert M.ntiles > 20 and len(set(geom[:, 6].astype(int))) >= 3
    lv = np.random.default_rng(3).choice([1, 2, 4, 8], size=M.ntiles).astype(np.int32)
    r = M.run(levels=lv)
    o = O.mesher_integrate(p, geom, lv)
    assert rel_err(r["m3mf"], o["m3mf"]) < 1e-12
    assert rel_err(r["m3mr"], o["m3mr"]) < 1e-12
    assert (r["evals_pre"], r["evals_post"]) == (o["evals_pre"], o["evals_post"])
    # rule-based levels when none are given
    a, b = M.run(), M.run(levels=rule)
    assert rel_err(a["m3mr"], b["m3mr"]) < 1e-14 and a["evals_post"] == b["evals_post"]
    M.close()


def test_level0_features_are_bfm_norms(gpu_lib):
    p = CASES["tiny_u"]()
    M = abi.Mesher(p, bkg_base_rfmnt=5, tile_cells=4, bkg_peakw=0.3)
    geom, _ = M.tiles()
    st = M.features(ns=2)                        # corners of each tile (one launch for all)
    for t in range(M.ntiles):
        f1 = np.array([geom[t, 0], geom[t, 1], geom[t, 0], geom[t, 1]])
        f2 = np.array([geom[t, 2], geom[t, 2], geom[t, 3], geom[t, 3]])
        want = np.linalg.norm(O.msh_points(p, f1, f2), axis=0)
        assert rel_err(st[t], want) < 1e-12
    M.close()


def test_refined_mesher_converges_to_classic(gpu_lib):
    """Tower of configuration 1 (f_n 0.2/0.6/1.2 Hz, 1 % damping), background
    width 0.1 Hz: every tile at level 8 against the Classic trapezoid integral
    with exact S(f1+f2) over the same square [-2.4, 2.4]^2 on a 0.25 mHz grid."""
    p = CASES["config1"]()
    M = abi.Mesher(p, bkg_peakw=0.1, diag_crest=True)
    r = M.run(levels=8)
    r1 = M.run(levels=1)
    M.close()
    M = abi.Mesher(p, bkg_peakw=0.1)
    r0 = M.run(levels=8)
    M.close()
    q = CASES["config1"]()
    q.nfreqs, q.df = 9601, 2.4 / 9600
    c = abi.classic_run(q, quad="trapezoid", exact=True)
    e8 = rel_err(r["m3mr"], c["m3mr"])
    e1 = rel_err(r1["m3mr"], c["m3mr"])
    e0 = rel_err(r0["m3mr"], c["m3mr"])
    print(f"Mesher vs Classic: level 8 {e8:.2e}, level 1 {e1:.2e}, level 8 without the "
          f"diagonal crest {e0:.2e}; BFM evals level 8 {r['evals_pre'] + r['evals_post']}, "
          f"BRM points {r['brm_points']}")
    # measured on MI355X: 2.3e-4 (diagonal crest), 1.5e-4 (without), level 1: 1.0e-2
    assert e8 < 1e-3 and e0 < 1e-3
    assert e1 > 3 * e8                           # BFM refinement matters
    assert rel_err(r["m3mf"], c["m3mf"]) < 1e-3


@pytest.mark.parametrize("name,tol", [("config1", 1e-3), ("config1", 1e-4), ("tiny_uvw", 1e-3)])
def test_adaptive_levels_meet_their_target(gpu_lib, name, tol):
    """bsahip_mesher_adapt: the a-posteriori levels keep m3mr within tol of the
    every-tile-level-8 result (relative to its largest entry), evaluating the
    BFM at no more points than level 8 everywhere (the lightly damped tower needs
    level 8 on nearly every tile); their count is the distinct points of those
    levels."""
    p = CASES[name]()
    M = abi.Mesher(p, bkg_base_rfmnt=20, tile_cells=8)
    conv = M.run(levels=8)
    lv, ev = M.adapt(tol)
    assert set(np.unique(lv)) <= {2, 4, 8}
    r = M.run(levels=lv)
    M.close()
    assert rel_err(r["m3mr"], conv["m3mr"]) <= tol
    assert ev == r["evals_pre"] + r["evals_post"]
    assert ev <= conv["evals_pre"] + conv["evals_post"]


@pytest.mark.parametrize("name,tol,diag", [("config1", 1e-3, False), ("tiny_uvw", 1e-3, False),
                                           ("tiny_uvw", 1e-4, False), ("tiny_dav", 1e-3, True)])
def test_adapt_run_equals_adapt_then_run(gpu_lib, name, tol, diag):
    """bsahip_mesher_adapt_run: the levels of bsahip_mesher_adapt and the moments
    and counts of bsahip_mesher_run at those levels (to the order of the sums),
    from the adaptation's own last integration."""
    p = CASES[name]()
    M = abi.Mesher(p, bkg_base_rfmnt=20, tile_cells=8, only_diag=diag)
    lv, ev = M.adapt(tol)
    want = M.run(levels=lv)
    got = M.adapt_run(tol)
    M.close()
    np.testing.assert_array_equal(got["levels"], lv)
    assert got["evals"] == ev
    for k in ("evals_pre", "evals_post", "brm_points"):
        assert got[k] == want[k], k
    assert rel_err(got["m3mf"], want["m3mf"]) < 1e-13
    assert rel_err(got["m3mr"], want["m3mr"]) < 1e-13
    print(f"{name} tol {tol}: levels {np.bincount(lv, minlength=9)[[2, 4, 8]]}")


def test_adapt_run_with_tiles_kept_below_level_8(gpu_lib):
    """Looser targets stop the refinement with tiles at levels 2 and 4 as well:
    adapt_run's moments then come from a run in which the kept tiles ran at
    their own levels.  At least one target must give mixed levels."""
    p = CASES["config1"]()
    M = abi.Mesher(p, bkg_base_rfmnt=20, tile_cells=8)
    mixed = 0
    for tol in (1e-1, 3e-2, 1e-2, 3e-3):
        lv, ev = M.adapt(tol)
        want = M.run(levels=lv)
        got = M.adapt_run(tol)
        np.testing.assert_array_equal(got["levels"], lv)
        assert got["evals"] == ev == want["evals_pre"] + want["evals_post"]
        assert rel_err(got["m3mf"], want["m3mf"]) < 1e-13
        assert rel_err(got["m3mr"], want["m3mr"]) < 1e-13
        mixed += np.unique(lv).size > 1
        print(f"tol {tol}: tiles at levels 2/4/8 {np.bincount(lv, minlength=9)[[2, 4, 8]]}")
    M.close()
    assert mixed > 0


def test_mesher_batch_equals_separate_runs(gpu_lib):
    """bsahip_mesher_run_batch over Meshers of different structures, mode counts
    and models (full and ONLY_DIAG), mixed and rule-based levels: each Mesher's
    moments equal its own bsahip_mesher_run (to the order of the atomic sums)
    and its counts are the same."""
    rng = np.random.default_rng(11)
    ms, lvs = [], []
    for name, kw in (("tiny_u", {}), ("tiny_uvw", {}), ("config1", {}), ("tiny_dav", {}),
                     ("tiny_u", {"only_diag": True})):
        M = abi.Mesher(CASES[name](), bkg_base_rfmnt=5, tile_cells=2, bkg_peakw=0.25, **kw)
        ms.append(M)
        lvs.append(rng.choice([1, 2, 4, 8], size=M.ntiles).astype(np.int32))
    lvs[2] = None                                 # rule-based levels for one of them
    sep = [M.run(levels=lv) for M, lv in zip(ms, lvs)]
    bat = abi.mesher_run_batch(ms, lvs)
    for a, b in zip(sep, bat):
        for k in ("m3mf", "m3mr"):
            assert rel_err(b[k], a[k]) < 1e-13, k
        assert (a["evals_pre"], a["evals_post"], a["brm_points"]) == \
            (b["evals_pre"], b["evals_post"], b["brm_points"])
    with pytest.raises(abi.BsaHipError):
        abi.mesher_run_batch([ms[0], ms[0]])
    for M in ms:
        M.close()
