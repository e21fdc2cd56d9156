"""CPU tests of the Mesher tile layout (bsahip_mesher_layout, host computation):
the 1D zone limits of prefetchZoneLimits_ (src/bsa/meshing/BsaMesherImpl.F90:1638-1859),
the peak extensions of getActualPeakZoneExtensionLimits_ (:1864-1890), the policy
BFM factors (policy/MPolicy.F90:130-168) and the delta validation of
zones/_shared_poly2d.fi:735-795, worked by hand for two modes."""
import ctypes as C

import numpy as np
import pytest

from bsalib_amd import abi, build


@pytest.fixture(scope="module", autouse=True)
def built():
    if build.needs_build():
        build.build(verbose=False)


def _cfg(natf, xsi, bkg_peakw=0.5, rfmnt=5, tile_cells=100, ext=2.0, diag=0):
    natf = np.ascontiguousarray(natf, dtype=np.float64)
    xsi = np.ascontiguousarray(xsi, dtype=np.float64)
    c = abi.MesherCfg()
    c.bkg_base_rfmnt, c.tile_cells, c.diag_crest = rfmnt, tile_cells, diag
    c.bkg_peakw, c.bkg_area_ext, c.peak_area_ext, c.max_area_ext = bkg_peakw, ext, ext, ext
    c.natf = natf.ctypes.data_as(C.POINTER(C.c_double))
    c.xsi = xsi.ctypes.data_as(C.POINTER(C.c_double))
    return c, (natf, xsi)


def test_two_mode_layout_by_hand():
    # b = 0.5*2/2 = 0.5; fmax = 3*2 = 6; peaks 1 +- 0.05*1*2 = [0.9, 1.1] (w 0.2),
    # 3 +- 0.02*3*2 = [2.88, 3.12] (w 0.24); df_ref = 2*0.5/(5-1) = 0.25
    c, keep = _cfg([3.0, 1.0], [0.02, 0.05])          # unsorted on purpose
    g, rule = abi.mesher_layout(c, 2)
    assert len(g) == 121
    ax = g[g[:, 2] == g[0, 2]]                         # first row of tiles: the f1 axis
    # BRM cells per level-0 cell: smallest power of 2 >= 8 with h0/sub <= max(w_n, d_n)/16
    # over the modes (d_n = distance of the interval to f_n)
    want = [  # a, b, cells, policy, BRM sub, rule level on this axis
        (-6.0, -3.12, 3, 9, 64, 2),     # PAD: h0 = 0.96 (4*df_ref), w = bkg 0.5, fct 1;
                                        #   BRM: d = 0.12 < w = 0.24 -> 0.96/(0.24/16) = 64
        (-3.12, -2.88, 1, 5, 16, 4),    # PEAK: h0 = 0.24, lim = 0.24/4; BRM 0.24/16 = w/16
        (-2.88, -1.1, 4, 4, 64, 2),     # PRE_PEAK_2: h0 = 0.445, lim = 0.5/2; BRM 0.445/0.0125
        (-1.1, -0.9, 1, 5, 16, 4),      # PEAK: h0 = 0.2, lim = 0.2/4
        (-0.9, -0.5, 1, 3, 32, 4),      # PRE_PEAK_1: h0 = 0.4, lim = 0.5/4; BRM 0.4/0.0125
        (-0.5, 0.5, 4, 100, 8, 2),      # background: h0 = 0.25, lim = 0.5/4; d = 0.5 -> 8
    ]
    want = want + [(-b, -a, n, p, s, l) for (a, b, n, p, s, l) in reversed(want[:-1])]
    assert len(ax) == len(want)
    lev = []
    for row, (a, b, n, pol, sub, l) in zip(ax, want):
        assert row[0] == pytest.approx(a, abs=1e-12) and row[1] == pytest.approx(b, abs=1e-12)
        assert (int(row[4]), int(row[6]), int(row[8])) == (n, pol, sub), (row, a, b)
        lev.append(l)
    lev = np.array(lev)
    np.testing.assert_array_equal(rule.reshape(11, 11), np.maximum(lev[:, None], lev[None, :]))


def test_peaks_inside_the_background_zone_give_one_zone():
    # b = 2.5*2/2 = 2.5 >= every peak and >= fmax = 1.2*2: a single background zone
    c, keep = _cfg([0.2, 0.6, 1.2], [0.01, 0.01, 0.01], bkg_peakw=2.5, rfmnt=9, tile_cells=8)
    g, rule = abi.mesher_layout(c, 3)
    assert len(g) == 1
    np.testing.assert_allclose(g[0, :4], [-2.5, 2.5, -2.5, 2.5])
    # the 0.2 Hz peak (w = 0.008) lies inside: BRM spacing 0.625/sub <= 0.0005 needs
    # sub = 2048, capped at 256
    assert (g[0, 4], g[0, 5], g[0, 6], g[0, 8]) == (8, 8, 100, 256)


def test_tiles_split_long_intervals_and_cover_the_domain():
    c, keep = _cfg([1.0, 3.0], [0.05, 0.02], bkg_peakw=0.5, rfmnt=21, tile_cells=3)
    g, _ = abi.mesher_layout(c, 2)
    assert g[:, 4].max() <= 3 and g[:, 5].max() <= 3
    area = np.sum((g[:, 1] - g[:, 0]) * (g[:, 3] - g[:, 2]))
    assert area == pytest.approx(12.0 ** 2, rel=1e-12)


def test_overlapping_peaks_merge():
    # peaks [0.9, 1.1] and [1.0, 1.2] overlap -> one PEAK interval [0.9, 1.2], w = 0.2
    c, keep = _cfg([1.0, 1.1], [0.05, 1.0 / 22.0], bkg_peakw=0.5, rfmnt=5)
    g, _ = abi.mesher_layout(c, 2)
    ax = g[g[:, 2] == g[0, 2]]
    peaks = ax[ax[:, 6] == 5]
    np.testing.assert_allclose(np.sort(peaks[:, 0]), [-1.2, 0.9], atol=1e-12)
    np.testing.assert_allclose(np.sort(peaks[:, 1]), [-0.9, 1.2], atol=1e-12)


def test_diagonal_crest_refines_tiles_crossing_f1_plus_f2_resonance():
    # background tile [-0.5,0.5]^2 (h0 = 0.25): f1+f2 spans [-1, 1] and crosses f = 1
    # (w = 0.2): with the diagonal crest its BRM spacing must reach 0.2/16
    for diag, sub in ((0, 8), (1, 32)):
        c, keep = _cfg([1.0, 3.0], [0.05, 0.02], diag=diag)
        g, _ = abi.mesher_layout(c, 2)
        bkg = g[(g[:, 6] == 100) & (g[:, 7] == 100)]
        assert len(bkg) == 1 and (int(bkg[0, 8]), int(bkg[0, 9])) == (sub, sub)


def _lattice():
    from tests.cases import CASES, LATTICE_MESHER
    p = CASES["tiny_uvw"]()
    kw = dict(LATTICE_MESHER)
    cfg, keep = abi.mesher_cfg(p, **kw)
    geom, _ = abi.mesher_layout(cfg, 2)
    return p, geom


def test_lattice_layout_tiles_share_one_brm_lattice():
    p, g = _lattice()
    assert len(g) == 256 and set(g[:, 6].astype(int)) == {3, 4, 5, 9, 100}
    # every tile: BRM spacing 1/16 on both axes, corners on the lattice k/16
    np.testing.assert_array_equal((g[:, 1] - g[:, 0]) / (g[:, 4] * g[:, 8]), 1.0 / 16.0)
    np.testing.assert_array_equal((g[:, 3] - g[:, 2]) / (g[:, 5] * g[:, 9]), 1.0 / 16.0)
    np.testing.assert_array_equal(np.round(g[:, :4] * 16.0), g[:, :4] * 16.0)
    np.testing.assert_array_equal(g[:, 8:10], 2)


def test_oracle_mesher_on_lattice_equals_classic_trapezoid():
    """Tile seams and edge weights against mathematics: BFM at every BRM point
    (level 2 = BRM cells per level-0 cell) on 256 tiles of five policies sharing one
    lattice, so the sum of the tiles' trapezoid rules (edges 1/2, vertices 1/4)
    is the tensor trapezoid on the union grid with S(f1+f2) evaluated exactly."""
    from oracle import oracle as O
    from tests.cases import CASES, LATTICE_CLASSIC
    from tests.conftest import rel_err
    p, g = _lattice()
    o = O.mesher_integrate(p, g, 2)
    q = CASES["tiny_uvw"]()
    q.nfreqs, q.df = LATTICE_CLASSIC["nfreqs"], LATTICE_CLASSIC["df"]
    c = O.classic_factorised(q, rule="trapezoid", truncate=False)
    assert rel_err(o["m3mf"], c["m3mf"]) < 1e-12
    assert rel_err(o["m3mr"], c["m3mr"]) < 1e-12
    assert o["evals_pre"] == 43 ** 2


def test_oracle_bilinear_interpolation_is_exact_on_bilinear_data():
    """Level 1 interpolates the BFM bilinearly onto 2 BRM cells per BFM cell;
    level 2 samples it at every BRM point.  On bilinear data both integrate the
    same values; a quadratic term makes them differ."""
    from oracle import oracle as O
    p, g = _lattice()
    rng = np.random.default_rng(7)
    c = rng.standard_normal((4, 8))

    def bilinear(prob, fi, fj, quad=0.0):
        fi, fj = np.asarray(fi), np.asarray(fj)
        return (c[0][:, None] + c[1][:, None] * fi + c[2][:, None] * fj
                + c[3][:, None] * fi * fj + quad * fi * fi)

    a = O.mesher_integrate(p, g, 1, points=bilinear)
    b = O.mesher_integrate(p, g, 2, points=bilinear)
    for k in ("m3mf", "m3mr"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-13, atol=1e-13 * np.abs(b[k]).max())
    q1 = O.mesher_integrate(p, g, 1, points=lambda *x: bilinear(*x, quad=1.0))
    q2 = O.mesher_integrate(p, g, 2, points=lambda *x: bilinear(*x, quad=1.0))
    assert np.abs(q1["m3mf"] - q2["m3mf"]).max() > 1e-6 * np.abs(q2["m3mf"]).max()
