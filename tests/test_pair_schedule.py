"""CPU check of the memory schedule of the paired-class 27-point sweep
(k_gs27_pair in multigrid_amd/csrc/pmg_kernels.hip): which array each row
update reads and writes, which rows each workgroup strip updates, and the
x <-> tmp ping-pong over a pair of sweeps.  The row update itself is the
Vgsrb27x point update (pmgc/gsd.cc:160-211); the schedule is run with the
workgroups in several orders and every order must reproduce the oracle's
8-colour sweep exactly, so a value read by one strip and written by another
in the same launch would show up as a mismatch."""
import numpy as np
import pytest

from tests import helpers as H


def count_parity(n, par):  # rows v in [2, n-1] with v % 2 == par
    first = 2 + par
    return 0 if first > n - 1 else (n - 1 - first) // 2 + 1


def row_update(n, ac, cc, f, j, k, pz, py, zmask, arrs, adj):
    """Update row (j, k) of class (pz, py) the way one wavefront does:
    both i colours in the sweep's order, neighbour rows from the arrays
    (own row, dy odd, dz odd, both odd), the new values into the last."""
    N, sy, sz = n ** 3, n, n * n
    pl = [ac[q * N:(q + 1) * N] for q in range(14)]
    oC, oE, oN, uC, oNE, oNW, uE, uW, uN, uS, uNE, uNW, uSE, uSW = pl

    def zero(dy, dz):
        return (zmask >> (2 * (pz ^ (dz != 0)) + (py ^ (dy != 0)))) & 1
    zN, zU, zNU = zero(1, 0), zero(0, 1), zero(1, 1)
    xO, xN, xU, xNU, xW = arrs
    rb = (j - 1) * n + (k - 1) * sz
    row = xO[rb:rb + n].copy()
    if (zmask >> (2 * pz + py)) & 1:
        row[1:n - 1] = 0.0

    def X(a, q, z):
        return 0.0 if z else a[q]

    def upd(i, xw, xe):
        p = rb + i - 1
        o = (oN[p] * X(xN, p + sy, zN) + oN[p - sy] * X(xN, p - sy, zN) + oE[p] * xe
             + oE[p - 1] * xw + oNE[p] * X(xN, p + 1 + sy, zN)
             + oNW[p] * X(xN, p - 1 + sy, zN) + oNW[p + 1 - sy] * X(xN, p + 1 - sy, zN)
             + oNE[p - 1 - sy] * X(xN, p - 1 - sy, zN))
        u = (uC[p] * X(xU, p + sz, zU) + uN[p] * X(xNU, p + sy + sz, zNU)
             + uS[p] * X(xNU, p - sy + sz, zNU) + uE[p] * X(xU, p + 1 + sz, zU)
             + uW[p] * X(xU, p - 1 + sz, zU) + uNE[p] * X(xNU, p + 1 + sy + sz, zNU)
             + uNW[p] * X(xNU, p - 1 + sy + sz, zNU) + uSE[p] * X(xNU, p + 1 - sy + sz, zNU)
             + uSW[p] * X(xNU, p - 1 - sy + sz, zNU))
        dd = (uC[p - sz] * X(xU, p - sz, zU) + uS[p + sy - sz] * X(xNU, p + sy - sz, zNU)
              + uN[p - sy - sz] * X(xNU, p - sy - sz, zNU) + uW[p + 1 - sz] * X(xU, p + 1 - sz, zU)
              + uE[p - 1 - sz] * X(xU, p - 1 - sz, zU)
              + uSW[p + 1 + sy - sz] * X(xNU, p + 1 + sy - sz, zNU)
              + uSE[p - 1 + sy - sz] * X(xNU, p - 1 + sy - sz, zNU)
              + uNW[p + 1 - sy - sz] * X(xNU, p + 1 - sy - sz, zNU)
              + uNE[p - 1 - sy - sz] * X(xNU, p - 1 - sy - sz, zNU))
        return (f[p] + ((o + u) + dd)) / (oC[p] + cc[p])
    for par in ([0, 1] if adj == 0 else [1, 0]):
        for i in range(2 + par, n, 2):
            row[i - 1] = upd(i, row[i - 2], row[i])
    xW[rb + 1:rb + n - 1] = row[1:n - 1]


def paired_sweeps(n, ac, cc, f, x0, nsweeps, adj, zero, rs, order):
    """nsweeps (even) sweeps as the host loop and kernel run them; `order`
    is the order in which the workgroups of a launch are executed."""
    x = x0.copy()
    tmp = x0.copy()  # same boundary values
    bp = 0 if adj else 1
    nstrips = -(-count_parity(n, bp) // rs)
    cur, nxt = x, tmp
    for sw in range(nsweeps):
        zmask = 0xF if (zero and sw == 0) else 0
        for h in range(2):
            pz = 1 - h if adj else h
            po = cur if h == 0 else nxt
            wgs = [(kk, st) for kk in range(count_parity(n, pz)) for st in range(nstrips)]

            def rows(st):
                jf = 2 + bp + 2 * rs * st
                jl = min(jf + 2 * (rs - 1), n - 1)
                first = [jf - 1 + 2 * w for w in range(rs + 1)
                         if 2 <= jf - 1 + 2 * w <= min(n - 1, jl + 1)]
                own = [jf + 2 * w for w in range(rs) if jf + 2 * w <= jl]
                return first, own

            def phase1(kk, st):
                for j in rows(st)[0]:
                    row_update(n, ac, cc, f, j, 2 + pz + 2 * kk, pz, 1 - bp, zmask,
                               (cur, cur, po, po, nxt), adj)

            def phase2(kk, st):
                for j in rows(st)[1]:
                    row_update(n, ac, cc, f, j, 2 + pz + 2 * kk, pz, bp,
                               zmask & ~(1 << (2 * pz + 1 - bp)), (cur, nxt, po, po, nxt), adj)
            if order == "first-phase-everywhere":
                for g in wgs:
                    phase1(*g)
                for g in wgs:
                    phase2(*g)
            else:
                seq = wgs if order == "forward" else list(reversed(wgs))
                if order == "shuffled":
                    seq = [wgs[i] for i in np.random.default_rng(len(wgs)).permutation(len(wgs))]
                for g in seq:
                    phase1(*g)
                    phase2(*g)
            zmask &= ~(0x3 << (2 * pz))
        cur, nxt = nxt, cur
    return x


@pytest.mark.parametrize("rs", [3, 7])
@pytest.mark.parametrize("adj,zero", [(0, False), (1, True)])
@pytest.mark.parametrize("order", ["first-phase-everywhere", "reverse", "shuffled"])
def test_paired_schedule_matches_oracle(orc, rs, adj, zero, order):
    n = 17  # 7 or 8 rows of a class per plane: several strips for rs = 3
    rng = np.random.default_rng(100 * rs + 10 * adj + zero)
    ac, cc = H.random_op(n, 27, rng)
    f = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    x0 = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    ref = orc.gsrb(n, 27, ac, cc, f, np.zeros_like(x0) if zero else x0, 2, adj)
    got = paired_sweeps(n, ac, cc, f, x0, 2, adj, zero, rs, order)
    assert np.array_equal(got, ref)

