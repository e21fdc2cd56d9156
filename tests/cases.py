"""Seeded problems used by the tests (same generators as the golden fixtures)."""
from bsalib_amd import synth

CASES = {
    "tiny_u": lambda: synth.tower(nn=6, nm=3, comps=(1,), nfreqs=9, df=0.1, seed=1),
    "tiny_uvw": lambda: synth.tower(nn=5, nm=2, comps=(1, 2, 3), nfreqs=7, df=0.15, seed=3),
    "tiny_dav": lambda: synth.tower(nn=5, nm=3, comps=(1,), nfreqs=8, df=0.12, seed=5, psd_id=5,
                                    zone_per_node=False),
    "config1": lambda: synth.config1(),
}
