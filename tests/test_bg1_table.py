"""The transcribed BG1 shift table (include/airan_ldpc_bg1.h) against a design
property of the standard's lifted graphs: at the largest lifting sizes of
set i_LS = 1 (Z = 96, 192, 384) the table has no 4-cycle -- no rows r1, r2
and columns c1, c2 with s(r1,c1) - s(r1,c2) + s(r2,c2) - s(r2,c1) = 0 mod Z.
A mistyped shift breaks this for about one single-digit typo in five
(measured by mutating the table), so it pins some transcription errors, not
all: the shift values themselves stay without an independent reference."""
import itertools
import re
from pathlib import Path

import pytest

HDR = Path(__file__).resolve().parents[1] / "include" / "airan_ldpc_bg1.h"


def _table():
    txt = HDR.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)

    def arr(name):
        m = re.search(name + r"\[[^\]]*\]\s*=\s*\{([^}]*)\}", txt)
        return [int(x) for x in re.findall(r"\d+", m.group(1))]

    rs, col, sh = arr("AIRAN_BG1_ROW_START"), arr("AIRAN_BG1_COL"), arr("AIRAN_BG1_SHIFT")
    assert len(rs) == 47 and len(col) == len(sh) == rs[-1] == 316
    return [[(col[e], sh[e]) for e in range(rs[r], rs[r + 1])] for r in range(46)]


def four_cycles(rows, z):
    n = 0
    for r1, r2 in itertools.combinations(range(len(rows)), 2):
        d1 = {c: s % z for c, s in rows[r1]}
        d2 = {c: s % z for c, s in rows[r2]}
        for a, b in itertools.combinations(sorted(set(d1) & set(d2)), 2):
            n += (d1[a] - d1[b] + d2[b] - d2[a]) % z == 0
    return n


@pytest.mark.parametrize("z", [96, 192, 384])
def test_bg1_lifted_graph_has_no_four_cycle(z):
    assert four_cycles(_table(), z) == 0


def test_four_cycle_check_catches_a_planted_typo():
    """The check is not vacuous: a shift chosen to close a 4-cycle at
    Z = 384 is found."""
    rows = _table()
    d0 = dict(rows[0])
    d1 = dict(rows[1])
    a, b = sorted(set(d0) & set(d1))[:2]
    k = [c for c, _ in rows[1]].index(b)
    rows[1][k] = (b, (d1[a] - d0[a] + d0[b]) % 384)  # s(1,b) := s(1,a) - s(0,a) + s(0,b)
    assert four_cycles(rows, 384) >= 1
