"""PSM TSV output (DFile_PrintScore, dslim_fileout.cpp:68-72,112-146)."""
import numpy as np

from gicops_amd import _ffi as F
from gicops_amd import tsv
from oracle import pyoracle as O


def test_header_matches_reference_columns():
    cols = tsv.HEADER.split("\t")
    assert cols == ["file", "scan_num", "prec_mass", "charge", "retention_time", "peptide", "matched_ions",
                    "total_ions", "calc_pep_mass", "mass_diff", "mod_info", "hyperscore", "expectscore",
                    "num_hits"]


def test_rows_for_reported_psms(small_index, small_batch, tmp_path):
    psm, _, _ = O.search(small_index, small_batch, O.default_params(dM=10.0))
    path = tmp_path / "out.tsv"
    n = tsv.write(path, small_index, small_batch, psm)
    lines = path.read_text().splitlines()
    assert lines[0] == tsv.HEADER
    assert n == len(lines) - 1 == int(((psm["flags"] & F.GX_PSM_REPORTED) != 0).sum())
    for line in lines[1:]:
        f = line.split("\t")
        assert len(f) == 14 and f[10] == ""
        s = int(f[1]) - 1
        r = psm[s]
        assert f[12] == "%f" % r["evalue"] and float(f[12]) < 20.0
        assert int(f[6]) == r["sharedions"] and int(f[13]) == r["cpsms"]
        b = small_index.buckets[r["idxoffset"]]
        assert f[5] == b.sequence(int(r["psid"])) and len(f[5]) == b.peplen
        # std::to_string prints six decimals
        assert all(len(x.split(".")[1]) == 6 for x in (f[2], f[4], f[8], f[9], f[11], f[12]))
