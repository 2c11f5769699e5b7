"""The C++ drop-in (reference-shaped DSLIM_QuerySpectrum over the C ABI) end
to end on the GPU: its TSV equals the TSV of the CPU oracle's results."""
import subprocess
import sys

import numpy as np
import pytest

from gicops_amd import _build, tsv
from gicops_amd import slm
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def test_cpp_dslim_queryspectrum_tsv(small_index, small_batch, tmp_path):
    print("case: C++ adapter end to end, small workload", flush=True)
    # the adapter_e2e program builds the same workload from the same seeds as the presets
    assert slm.DB_PRESETS["small"].n_peptides == 50_000 and slm.DB_PRESETS["small"].seed == 0
    sc = slm.SPEC_PRESETS["small"]
    assert (sc.n_spectra, sc.seed, sc.noise_min, sc.noise_max, sc.qalen) == (1000, 0, 20, 80, 100)
    exe = _build.LIBDIR / "adapter_e2e"
    out = tmp_path / "gpu.tsv"
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    psm, _, _ = O.search(small_index, small_batch, O.default_params(dM=10.0))
    ref = tmp_path / "cpu.tsv"
    tsv.write(ref, small_index, small_batch, psm)
    got = out.read_text().splitlines()
    exp = ref.read_text().splitlines()
    assert got[0] == exp[0] == tsv.HEADER
    assert len(got) == len(exp) and len(exp) > 1
    for g, e in zip(sorted(got[1:]), sorted(exp[1:])):
        gf, ef = g.split("\t"), e.split("\t")
        assert gf[:12] == ef[:12] and gf[13] == ef[13], (g, e)
        assert abs(float(gf[12]) - float(ef[12])) <= 1e-6 * max(1.0, abs(float(ef[12]))) + 1e-6
