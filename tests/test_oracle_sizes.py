"""Sizing arithmetic of Vmgsz / Vmaxlev / Vbuildstr (pmgc/mgdrvd.cc:505-601,
pmgc/mgsubd.cc:266-347, 456-484) against closed forms."""
import numpy as np
import pytest


def levels(n, nlev):
    out = [n]
    for _ in range(nlev - 1):
        out.append((out[-1] - 1) // 2 + 1)
    return out


@pytest.mark.parametrize("n,nlev", [(65, 4), (161, 5), (289, 6), (129, 5), (97, 5), (17, 3)])
def test_mgsz_closed_form(orc, n, nlev):
    rc, out = orc.mgsz(2, 0, 0, n, n, n, nlev)
    assert rc == 0
    lv = levels(n, nlev)
    nf = n ** 3
    narr = sum(m ** 3 for m in lv)
    narrc = narr - nf
    assert list(out[:3]) == [lv[-1]] * 3
    assert out[3] == nf and out[5] == narr and out[6] == narrc
    assert out[10] == 2 * narr + 4 * nf + 41 * narrc + 100 * (nlev + 1)
    assert out[11] == 150 * (nlev + 1)


def test_mgsz_workload_numbers(orc):
    """Sizes quoted for the 65^3 / 161^3 / 289^3 workloads."""
    _, o = orc.mgsz(2, 0, 0, 65, 65, 65, 4)
    assert (o[3], o[5], o[6], o[10]) == (274625, 316204, 41579, 3436147)
    _, o = orc.mgsz(2, 0, 0, 161, 161, 161, 5)
    assert (o[3], o[5]) == (4173281, 4784235)
    _, o = orc.mgsz(2, 0, 0, 289, 289, 289, 6)
    assert (o[3], o[5]) == (24137569, 27633723)


@pytest.mark.parametrize("n,expect", [(65, 6), (161, 6), (289, 6), (129, 7), (97, 6), (9, 3), (10, 1)])
def test_maxlev(orc, n, expect):
    # closed form: largest L with (n-1) divisible by 2^(L-1) and coarse size > 2
    L = 0
    while True:
        iden = 2 ** L
        if (n - 1) % iden or (n - 1) // iden + 1 <= 2:
            break
        L += 1
    assert L == expect
    assert orc.olib().pmgo_maxlev(n, n, n) == expect


def test_buildstr_offsets(orc):
    n, nlev = 33, 3
    iz = np.zeros(50 * (nlev + 1), np.int32)
    orc.olib().pmgo_buildstr(n, n, n, nlev, orc.ip(iz))
    IZ = lambda a, lev: iz[(lev - 1) * 50 + a - 1]
    lv = levels(n, nlev)
    assert IZ(1, 1) == 1 and IZ(1, 2) == 1 + lv[0] ** 3 and IZ(1, 3) == 1 + lv[0] ** 3 + lv[1] ** 3
    assert IZ(11, 2) == 1 + 27 * lv[1] ** 3
    assert IZ(5, 3) == 201 and IZ(8, 2) == 1 + n
