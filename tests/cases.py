"""Seeded problems used by the tests (same generators as the golden fixtures)."""
from bsalib_amd import synth

CASES = {
    "tiny_u": lambda: synth.tower(nn=6, nm=3, comps=(1,), nfreqs=9, df=0.1, seed=1),
    "tiny_uvw": lambda: synth.tower(nn=5, nm=2, comps=(1, 2, 3), nfreqs=7, df=0.15, seed=3),
    "tiny_dav": lambda: synth.tower(nn=5, nm=3, comps=(1,), nfreqs=8, df=0.12, seed=5, psd_id=5,
                                    zone_per_node=False),
    "config1": lambda: synth.config1(),
    # three DOFs per node, DOFs 1 and 2 loaded (NLIBSL = 2; the reference's
    # Mesher BFM reads wfc by DOF number, so loaded DOFs must lead: trap T21)
    "tiny_2dof": lambda: synth.tower(nn=5, nm=3, comps=(1, 3), nfreqs=8, df=0.125, seed=7,
                                     nlibs=3, libs_load=(1, 2)),
}


# Mesher layout on one uniform lattice (used with CASES["tiny_uvw"]): b = 0.5,
# df_ref = 2b/(9-1) = 1/8; peaks 1 +- 0.1875*1*2 = [0.625, 1.375] and
# 2 +- 0.125*2*2 = [1.5, 2.5]; pre-peak gaps [0.5, 0.625], [1.375, 1.5] and the
# padding [2.5, 2*1.3125] are one df_ref long, so every level-0 cell is df_ref
# and, with 2 BRM cells per level-0 cell, every tile's BRM grid lies on the
# lattice k/16 over [-2.625, 2.625]^2 (the Classic grid nfreqs = 43, df = 1/16).
LATTICE_MESHER = dict(bkg_base_rfmnt=9, tile_cells=4, bkg_peakw=0.5, bkg_area_ext=2.0,
                      peak_area_ext=2.0, max_area_ext=1.3125, natf=[1.0, 2.0],
                      xsi=[0.1875, 0.125], brm_cells=2)
LATTICE_CLASSIC = dict(nfreqs=43, df=1.0 / 16.0)
