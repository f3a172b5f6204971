"""The C-ABI library loads, exports every declared symbol, and its host-only
entry points (sizing, synthetic inputs) work without a GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT, have_gpu


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "pmg_mi355x.h")).read()
    return sorted(set(re.findall(r"\b(pmg_\w+)\s*\(", src)))


def test_library_exports_all_declared(mg):
    out = subprocess.run(["nm", "-D", "--defined-only", mg.LIBPATH],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (pmg_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert len(declared_symbols()) >= 25


def test_mgsz_matches_oracle(mg, orc):
    for n, nlev in [(65, 4), (161, 5), (289, 6), (33, 3)]:
        _, o = orc.mgsz(2, 0, 0, n, n, n, nlev)
        s = mg.Vmgsz(2, 0, 0, n, n, n, nlev)
        assert [s[k] for k in ("nxc", "nyc", "nzc", "nf", "nc", "narr", "narrc", "n_rpc",
                               "n_iz", "n_ipc", "iretot", "iintot")] == list(o)
        assert mg.Vmaxlev(n, n, n) == orc.olib().pmgo_maxlev(n, n, n)


def test_bad_sizes_rejected(mg):
    with pytest.raises(mg.PMGError):
        mg.Vmgsz(2, 5, 0, 33, 33, 33, 3)


def test_packmg_layout(mg):
    iparm, rparm = mg.Vpackmg(10, 20, 33, 33, 33, 3, errtol=1e-7)
    assert iparm.shape == (22,) and list(iparm[:6]) == [10, 20, 33, 33, 33, 3]
    assert iparm[16] == 1 and iparm[17] == 2 and rparm[0] == 1e-7


@pytest.mark.skipif(have_gpu(), reason="checks the no-GPU error path")
def test_no_gpu_is_reported(mg, orc):
    n, nlev = 9, 2
    inp = orc.pbe_synth(n, 0.5, 4, 1.0)
    with pytest.raises(mg.PMGError) as e:
        mg.pbe_assemble(n, 0.5, 0.15, inp["atoms"])
    assert e.value.code == mg.PMG_EDEVICE
    iparm, rparm = mg.make_params(n, nlev)
    sz = mg.Vmgsz(2, 0, 0, n, n, n, nlev)
    with pytest.raises(mg.PMGError) as e:
        mg.Vmgdriv(iparm, rparm, np.zeros(sz["iintot"], np.int32), np.zeros(sz["iretot"]),
                   np.zeros(n ** 3), *(inp[k] for k in ("xf", "yf", "zf", "gxcf", "gycf",
                                                        "gzcf", "a1cf", "a2cf", "a3cf",
                                                        "ccf", "fcf", "tcf")))
    assert e.value.code == mg.PMG_EDEVICE


@pytest.mark.skipif(have_gpu(), reason="checks the no-GPU error path")
def test_server_without_gpu_is_reported(mg):
    with pytest.raises(mg.PMGError) as e:
        mg.Server(0, 2)
    assert e.value.code == mg.PMG_EDEVICE
