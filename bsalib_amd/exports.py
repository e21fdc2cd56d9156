"""Reader of the bispectrum export file (bsaexport_cls.brm and its kin).

Layout, as the reference writes it (exportBRM_baseHeaderWriter_internal_ and
exportBRM_base_internal_, src/bsa/BsaLibImpl.F90:1849-1913, with the point
count patched in by writeNTotPtsToBispUn_, :1832-1845), an unformatted stream:

* int32 nm, then nm int32 indexes (3 modal indexes, or node and DOF);
* int32 combinations, spatial symmetry, number of zones, total points;
* per zone: int32 zone id (1 = rectangle), nI, nJ, then nI*nJ points of
  real32 fi, real32 fj and `nval` real32 values (1 in visual mode).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def read_bisp_export(path, nval: int = 1) -> dict:
    buf = Path(path).read_bytes()
    pos = 0

    def ints(n):
        nonlocal pos
        v = np.frombuffer(buf, dtype=np.int32, count=n, offset=pos)
        pos += 4 * n
        return v.copy()

    nm = int(ints(1)[0])
    out = {"nm": nm, "indexes": ints(nm)}
    out["ncomb"], out["ispsym"], out["nzones"], out["npts"] = (int(v) for v in ints(4))
    zones = []
    for _ in range(out["nzones"]):
        zid, ni, nj = (int(v) for v in ints(3))
        n = ni * nj
        rec = np.frombuffer(buf, dtype=np.float32, count=n * (2 + nval), offset=pos)
        pos += 4 * n * (2 + nval)
        rec = rec.reshape(n, 2 + nval)
        zones.append({"id": zid, "ni": ni, "nj": nj, "fi": rec[:, 0].copy(),
                      "fj": rec[:, 1].copy(), "values": rec[:, 2:].copy()})
    if pos != len(buf):
        raise ValueError(f"{path}: {len(buf) - pos} trailing bytes")
    out["zones"] = zones
    return out
