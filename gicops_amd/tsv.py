"""PSM output in the reference's TSV format.

Columns and formatting follow DFile_InitFiles / DFile_PrintScore
(source/core/dslim_fileout.cpp:68-72,112-146): one row per spectrum whose
e-value passes the report gate (e < expect_max, dslim_query.cpp:1045), values
rendered like std::to_string (integers plain, float and double as "%f").
"""
from __future__ import annotations

import numpy as np

from . import _ffi as F

HEADER = ("file\tscan_num\tprec_mass\tcharge\tretention_time\tpeptide\tmatched_ions\ttotal_ions\t"
          "calc_pep_mass\tmass_diff\tmod_info\thyperscore\texpectscore\tnum_hits")


def _f(x) -> str:
    # std::to_string(float/double) == printf("%f")
    return "%f" % float(x)


def rows(index, batch, psms: np.ndarray, file_name: str = "query.ms2", first_spec_id: int = 0):
    """Yield the TSV rows DFile_PrintScore would print for a searched batch."""
    rep = np.nonzero((psms["flags"] & F.GX_PSM_REPORTED) != 0)[0]
    for s in rep:
        r = psms[s]
        b = index.buckets[int(r["idxoffset"])]
        psid = int(r["psid"])
        pmass = np.float32(batch.precurse[s])
        mass = np.float32(b.mass[psid])
        yield "\t".join([
            file_name,
            str(first_spec_id + int(s) + 1),
            _f(pmass),
            str(int(batch.charges[s])),
            _f(np.float32(batch.rtimes[s])),
            b.sequence(psid),
            str(int(r["sharedions"])),
            str(int(r["totalions"])),
            _f(mass),
            _f(np.float32(pmass - mass)),
            "",
            _f(r["hyperscore"]),
            _f(r["evalue"]),
            str(int(r["cpsms"])),
        ])


def write(path, index, batch, psms, file_name: str = "query.ms2", first_spec_id: int = 0) -> int:
    n = 0
    with open(path, "w") as f:
        f.write(HEADER + "\n")
        for line in rows(index, batch, psms, file_name, first_spec_id):
            f.write(line + "\n")
            n += 1
    return n
