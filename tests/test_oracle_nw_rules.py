"""Hand-derived pins for the oracle's Needleman-Wunsch distance.

`oracle_nw_pair` restates genDist's CPU path: the banded form
`computeGenDist_CPU_band` (genCPU.cpp:348-464) for band >= 2 and the full
form `computeGenDist_CPU_full` (genCPU.cpp:132-225) for band <= 1.  No
library routine has these penalties and tie rules, so each case below is
small enough to be stepped through cell by cell, and its docstring is that
walk.  The expected (ML, AL, score) were derived by hand from the reference
lines quoted; the test only compares the oracle with them.

Notation (genCPU.cpp:356-384):
  s1 = read[pair[0]], length L1, the inner loop over columns j;
  s2 = read[pair[1]], length L2, the outer loop over rows i;
  kB = (L1 + L2) / (2 band)                                   (:365)
  a cell (i, j) is computed only when |i - j| <= kB           (insideBand, :76-79)
Boundary (:367-379, :390-396): (0,0) and, for 1 <= t <= kB, (0,t) and (t,0)
hold M 0, U 0, ML 0, AL t.  In the full form the whole of row 0 and
column 0 is set that way (:147-161, :172-175).
Scores: match +5, mismatch -4, raw bytes compared (:403-411, genMain.h:34-35);
GO = -10, GE = -1 (genMain.h:32-33).
Moves into (i, j) (:413-423):
  diag = M(i-1,j-1) + score
  up   = M(i,j-1) + (U&1)*GO + (U>>2&1)*GE   with U = U(i,j-1); -1000000 if (i,j-1) is out of band
  left = M(i-1,j) + (U&1)*GO + (U>>1&1)*GE   with U = U(i-1,j); -1000000 if (i-1,j) is out of band
So a gap opened after a diagonal costs GO, the same direction again costs GE,
and a switch of direction (up after left, U = 2; left after up, U = 4) costs
nothing; a gap leaving the boundary (U = 0) costs nothing either.
M = max(diag, up, left); the first of diag, up, left that equals M wins
(:425-440): diag -> U 1, ML = ML(i-1,j-1) - match + 1, AL = AL(i-1,j-1) + 1;
up -> U 4, ML/AL of (i,j-1) + 1; left -> U 2, ML/AL of (i-1,j) + 1.
End cells (:442-454): over row i = L2 keep the last cell with M >= the best so
far (bestCol, the reference's maxMLastCol), over column j = L1 likewise
(bestRow, maxMLastRow); both start at -1000000.  Result (:461-464): bestCol >
bestRow ? MLCol/ALCol : MLRow/ALRow, as float.

Tables list in-band cells row by row as
  (i,j) s2[i-1]s1[j-1]: diag/up/left -> M, winner, ML/AL
with "x" for -1000000.

`test_hand_cases_kill_mutations` plants plausible mistakes in a copy of the
oracle source (on the CPU, built with gcc into a temporary directory) and
checks that the case named for each one tells it apart.
"""
import ctypes
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "oracle", "crispy_oracle.c")

# name -> (s1, s2, band, ML, AL, score): filled in by the case functions below
CASES = {}


def case(name, s1, s2, band, ml, al, score):
    CASES[name] = (s1.encode(), s2.encode(), band, ml, al, score)

    def check():
        d, got_ml, got_al, got_sc = O.nw_pair(s1.encode(), s2.encode(), band)
        assert (got_ml, got_al, got_sc) == (ml, al, score), (name, got_ml, got_al, got_sc)
        assert np.float32(d).view(np.uint32) == np.float32(np.float32(ml) / np.float32(al)).view(np.uint32)
    return check


def test_gap_leaving_boundary_then_free_switch():
    """s1 = "A", s2 = "CCC", band 2: kB = 4/4 = 1.  Zero-cost switch
    left-after-up (:418-421 with U = 4: 4&1 = 0, 4>>1&1 = 0) and the band
    edge on "up" (:413-416).

      (1,1) CA: diag M(0,0)-4 = -4; up from boundary (1,0): 0 + 0 = 0;
                left from boundary (0,1): 0 -> M 0, diag no, up yes -> U 4,
                ML/AL = (1,0)'s 0/1 + 1 = 1/2
      (2,1) CA: diag M(1,0)-4 = -4; up: (2,0) out of band (|2-0| > 1) -> x;
                left from (1,1) with U 4: 0 + 0 + 0 = 0 -> M 0, left, U 2,
                ML/AL = 1/2 + 1 = 2/3
      (3,1): |3-1| > 1, out of band.
    Row i = L2 = 3 has no in-band cell: bestCol stays -1000000.
    Column j = L1 = 1: (1,1) M 0 -> bestRow 0 (1/2); (2,1) M 0 >= 0 ->
    bestRow 0 (2/3).  -1000000 > 0 is false -> bestRow's cell: 2/3, score 0.
    A left that charged GE after an up would make (2,1) = -1 (still left,
    2/3) and the first bestRow cell (1,1) would win: 1/2."""
    case("H1", "A", "CCC", 2, 2, 3, 0)()


def test_free_switch_up_after_left():
    """s1 = "AA", s2 = "CC", band 2: kB = 1.  Up-after-left is free
    (:413-416 with U = 2: 2&1 = 0, 2>>2&1 = 0).

      (1,1) CA: diag -4, up 0 (boundary), left 0 (boundary) -> 0, up, 1/2
      (1,2) CA: diag M(0,1)-4 = -4; up from (1,1) U 4: 0 + GE = -1;
                left: (0,2) out of band -> x -> -1, up, 2/3
      (2,1) CA: diag M(1,0)-4 = -4; up: (2,0) out of band -> x;
                left from (1,1) U 4: 0 -> 0, left, 2/3
      (2,2) CA: diag M(1,1)-4 = -4; up from (2,1) U 2: 0 + 0 + 0 = 0;
                left from (1,2) U 4: -1 + 0 = -1 -> 0, up, ML/AL of (2,1)
                + 1 = 3/4
    Row 2: (2,1) 0 -> bestCol 0 (2/3); (2,2) 0 >= 0 -> 0 (3/4).
    Column 2: (1,2) -1 -> bestRow -1 (2/3); (2,2) 0 -> 0 (3/4).
    0 > 0 false -> bestRow's cell: 3/4, score 0.  If the up move after a left paid
    GE, (2,2) = max(-4, -1, -1) = -1: row 2 keeps (2,1) (0, 2/3) and
    bestCol 0 > bestRow -1 takes it: 2/3."""
    case("H2", "AA", "CC", 2, 3, 4, 0)()


def test_last_row_keeps_last_maximum():
    """s1 = "AAA", s2 = "CA", band 2: kB = 5/4 = 1.  Row i = L2 keeps the
    last cell with M >= best (:442-447).

      (1,1) CA: -4 / 0 / 0 -> 0, up, 1/2
      (1,2) CA: diag M(0,1)-4 = -4; up (1,1) U 4: 0 - 1 = -1; left: (0,2)
                out -> x -> -1, up, 2/3
      (2,1) AA: diag M(1,0)+5 = 5; up: (2,0) out -> x; left (1,1) U 4: 0
                -> 5, diag, ML = 0 - 1 + 1 = 0, AL = 1 + 1 = 2 -> 0/2
      (2,2) AA: diag M(1,1)+5 = 5; up (2,1) U 1: 5 - 10 = -5; left (1,2)
                U 4: -1 -> 5, diag, ML = 1 - 1 + 1 = 1, AL = 2 + 1 = 3 -> 1/3
      (2,3) AA: diag M(1,2)+5 = 4; up (2,2) U 1: -5; left: (1,3) out -> x
                -> 4, diag, ML = 2 - 1 + 1 = 2, AL = 4 -> 2/4
    Row 2: (2,1) 5 -> bestCol 5 (0/2); (2,2) 5 >= 5 -> (1/3); (2,3) 4 no.
    Column 3: only (2,3) in band -> bestRow 4 (2/4).
    5 > 4 -> bestCol's cell: 1/3, score 5.  Keeping the first maximum gives 0/2."""
    case("H3", "AAA", "CA", 2, 1, 3, 5)()


def test_final_tie_prefers_last_column():
    """s1 = "AAAC", s2 = "ACAA", band 2: kB = 8/4 = 2.  bestCol == bestRow
    takes the column j = L1 cell (:461-464: only a strictly greater bestCol
    picks the row i = L2 cell).

      (1,1) AA: 5 / 0 / 0 -> 5, diag, 0/1
      (1,2) AA: diag M(0,1)+5 = 5; up (1,1) U 1: -5; left (0,2) boundary 0
                -> 5, diag, ML 0 - 1 + 1 = 0, AL 1 + 1 = 2 -> 0/2
      (1,3) AA: diag M(0,2)+5 = 5; up -5; left (0,3) out -> x -> 5, diag, 0/3
      (2,1) CA: diag M(1,0)-4 = -4; up (2,0) boundary: 0; left (1,1) U 1:
                -5 -> 0, up, ML/AL (2,0) 0/2 + 1 = 1/3
      (2,2) CA: diag M(1,1)-4 = 1; up (2,1) U 4: -1; left (1,2) U 1: -5
                -> 1, diag, 0 + 1 = 1 / 1 + 1 = 2 -> 1/2
      (2,3) CA: diag M(1,2)-4 = 1; up (2,2) U 1: -9; left (1,3): -5
                -> 1, diag, 1/3
      (2,4) CC: diag M(1,3)+5 = 10; up -9; left (1,4) out -> x
                -> 10, diag, 0/4
      (3,1) AA: diag M(2,0)+5 = 5; up (3,0) out -> x; left (2,1) U 4: 0
                -> 5, diag, 0/3
      (3,2) AA: diag M(2,1)+5 = 5; up (3,1) U 1: -5; left (2,2) U 1: -9
                -> 5, diag, 1/4
      (3,3) AA: diag M(2,2)+5 = 6; up -5; left -9 -> 6, diag, 1/3
      (3,4) AC: diag M(2,3)-4 = -3; up (3,3) U 1: -4; left (2,4) U 1: 0
                -> 0, left, 0/4 + 1 = 1/5
      (4,2) AA: diag M(3,1)+5 = 10; up (4,1) out -> x; left (3,2) U 1: -5
                -> 10, diag, 0/4
      (4,3) AA: diag M(3,2)+5 = 10; up (4,2) U 1: 0; left (3,3) U 1: -4
                -> 10, diag, 1/5
      (4,4) AC: diag M(3,3)-4 = 2; up (4,3) U 1: 0; left (3,4) U 2: 0 - 1
                = -1 -> 2, diag, 2/4
    Row 4: (4,2) 10 -> 0/4; (4,3) 10 >= 10 -> bestCol 10 (1/5); (4,4) 2.
    Column 4: (2,4) 10 -> bestRow 10 (0/4); (3,4) 0; (4,4) 2.
    10 > 10 false -> bestRow's cell: 0/4, score 10 (bestCol's would give 1/5)."""
    case("H4", "AAAC", "ACAA", 2, 0, 4, 10)()


def test_diag_wins_a_tie_with_up():
    """s1 = "ACAG", s2 = "CTCGC", band 2: kB = 9/4 = 2.  diag is tested
    before up (:425-435), so a diag/up tie is a diagonal.

      (1,1) CA: -4 / 0 / 0 -> 0, up, 1/2
      (1,2) CC: diag M(0,1)+5 = 5; up (1,1) U 4: -1; left (0,2): 0 -> 5, diag, 0/2
      (1,3) CA: diag M(0,2)-4 = -4; up (1,2) U 1: -5; left (0,3) out -> x
                -> -4, diag, 0 + 1 = 1 / 2 + 1 = 3 -> 1/3
      (2,1) TA: diag M(1,0)-4 = -4; up (2,0): 0; left (1,1) U 4: 0 -> 0, up, 1/3
      (2,2) TC: diag M(1,1)-4 = -4; up (2,1) U 4: -1; left (1,2) U 1: -5
                -> -1, up, 2/4
      (2,3) TA: diag M(1,2)-4 = 1; up (2,2) U 4: -2; left (1,3) U 1: -14
                -> 1, diag, 1/3
      (2,4) TG: diag M(1,3)-4 = -8; up (2,3) U 1: -9; left (1,4) out -> x
                -> -8, diag, 2/4
      (3,1) CA: diag M(2,0)-4 = -4; up (3,0) out -> x; left (2,1) U 4: 0
                -> 0, left, 2/4
      (3,2) CC: diag M(2,1)+5 = 5; up (3,1) U 2: 0; left (2,2) U 4: -1
                -> 5, diag, 1/4
      (3,3) CA: diag M(2,2)-4 = -5; up (3,2) U 1: 5 - 10 = -5; left (2,3)
                U 1: -9 -> -5: diag == up, diag wins, ML 2 + 1 = 3,
                AL 4 + 1 = 5 -> 3/5 (up would give 2/5)
      (3,4) CG: diag M(2,3)-4 = -3; up (3,3) U 1: -15; left (2,4): -18
                -> -3, diag, 2/4
      (4,2) GC: diag M(3,1)-4 = -4; up (4,1) out -> x; left (3,2) U 1: -5
                -> -4, diag, 3/5
      (4,3) GA: diag M(3,2)-4 = 1; up -14; left (3,3) U 1: -15 -> 1, diag, 2/5
      (4,4) GG: diag M(3,3)+5 = 0; up (4,3) U 1: -9; left (3,4) U 1: -13
                -> 0, diag, ML 3 - 1 + 1 = 3, AL 6 -> 3/6
      (5,3) CA: diag M(4,2)-4 = -8; up (5,2) out -> x; left (4,3) U 1: -9
                -> -8, diag, 4/6
      (5,4) CG: diag M(4,3)-4 = -3; up (5,3) U 1: -18; left (4,4) U 1: -10
                -> -3, diag, 3/6
    Row 5: (5,3) -8, (5,4) -3 -> bestCol -3 (3/6).
    Column 4: (2,4) -8, (3,4) -3, (4,4) 0, (5,4) -3 -> bestRow 0 (3/6).
    -3 > 0 false -> bestRow's cell: 3/6, score 0.  With up winning the tie at
    (3,3), (4,4) inherits 2/5 + 1 = 2/6."""
    case("H5", "ACAG", "CTCGC", 2, 3, 6, 0)()


def test_diag_wins_a_tie_with_left():
    """s1 = "CGCAC", s2 = "ACTA", band 2: kB = 9/4 = 2.  diag is tested
    before left (:425-440), so a diag/left tie is a diagonal.

      (1,1) AC: -4 / 0 / 0 -> 0, up, 1/2
      (1,2) AG: diag M(0,1)-4 = -4; up (1,1) U 4: -1; left (0,2): 0 -> 0, left, 1/3
      (1,3) AC: diag M(0,2)-4 = -4; up (1,2) U 2: 0; left (0,3) out -> x
                -> 0, up, 2/4
      (2,1) CC: diag M(1,0)+5 = 5; up (2,0): 0; left (1,1) U 4: 0 -> 5, diag, 0/2
      (2,2) CG: diag M(1,1)-4 = -4; up (2,1) U 1: -5; left (1,2) U 2: -1
                -> -1, left, 2/4
      (2,3) CC: diag M(1,2)+5 = 5; up (2,2) U 2: -1; left (1,3) U 4: 0
                -> 5, diag, ML 1 - 1 + 1 = 1, AL 4 -> 1/4
      (2,4) CA: diag M(1,3)-4 = -4; up (2,3) U 1: -5; left (1,4) out -> x
                -> -4, diag, 3/5
      (3,1) TC: diag M(2,0)-4 = -4; up (3,0) out -> x; left (2,1) U 1: -5
                -> -4, diag, 1/3
      (3,2) TG: diag M(2,1)-4 = 1; up (3,1) U 1: -14; left (2,2) U 2: -2
                -> 1, diag, 1/3
      (3,3) TC: diag M(2,2)-4 = -5; up (3,2) U 1: -9; left (2,3) U 1: -5
                -> -5: diag == left, diag wins, ML 2 + 1 = 3, AL 4 + 1 = 5
                -> 3/5 (left would give 2/5)
      (3,4) TA: diag M(2,3)-4 = 1; up (3,3) U 1: -15; left (2,4) U 1: -14
                -> 1, diag, 2/5
      (3,5) TC: diag M(2,4)-4 = -8; up (3,4) U 1: -9; left (2,5) out -> x
                -> -8, diag, 4/6
      (4,2) AG: diag M(3,1)-4 = -8; up (4,1) out -> x; left (3,2) U 1: -9
                -> -8, diag, 2/4
      (4,3) AC: diag M(3,2)-4 = -3; up (4,2) U 1: -18; left (3,3) U 1: -15
                -> -3, diag, 2/4
      (4,4) AA: diag M(3,3)+5 = 0; up (4,3) U 1: -13; left (3,4) U 1: -9
                -> 0, diag, ML 3 - 1 + 1 = 3, AL 6 -> 3/6
      (4,5) AC: diag M(3,4)-4 = -3; up (4,4) U 1: -10; left (3,5) U 1: -18
                -> -3, diag, 3/6
    Row 4: (4,2) -8, (4,3) -3, (4,4) 0, (4,5) -3 -> bestCol 0 (3/6).
    Column 5: (3,5) -8, (4,5) -3 -> bestRow -3 (3/6).
    0 > -3 -> bestCol's cell: 3/6, score 0.  With left winning the tie at
    (3,3), (4,4) inherits 2/5 and ends at 2/6."""
    case("H8", "CGCAC", "ACTA", 2, 3, 6, 0)()


def test_gap_open_after_diagonal_costs_go_only():
    """s1 = "ACAG", s2 = "ACGA", band 2: kB = 2.  The first gap after a
    diagonal costs GO = -10 and no GE (U = 1: 1>>2&1 = 0, 1>>1&1 = 0).

      (1,1) AA: 5 / 0 / 0 -> 5, diag, 0/1
      (1,2) AC: diag M(0,1)-4 = -4; up (1,1) U 1: 5 - 10 = -5; left (0,2): 0
                -> 0, left, ML/AL (0,2) 0/2 + 1 = 1/3
      (1,3) AA: diag M(0,2)+5 = 5; up (1,2) U 2: 0; left (0,3) out -> x
                -> 5, diag, 0/3
      (2,1) CA: diag M(1,0)-4 = -4; up (2,0): 0; left (1,1) U 1: -5
                -> 0, up, 1/3
      (2,2) CC: diag M(1,1)+5 = 10; up (2,1) U 4: -1; left (1,2) U 2: -1
                -> 10, diag, 0/2
      (2,3) CA: diag M(1,2)-4 = -4; up (2,2) U 1: 0; left (1,3) U 1: -5
                -> 0, up, 1/3
      (2,4) CG: diag M(1,3)-4 = 1; up (2,3) U 4: -1; left (1,4) out -> x
                -> 1, diag, 1/4
      (3,1) GA: diag M(2,0)-4 = -4; up (3,0) out -> x; left (2,1) U 4: 0
                -> 0, left, 2/4
      (3,2) GC: diag M(2,1)-4 = -4; up (3,1) U 2: 0; left (2,2) U 1: 0
                -> 0, up (tested before left), 3/5
      (3,3) GA: diag M(2,2)-4 = 6; up (3,2) U 4: -1; left (2,3) U 4: 0
                -> 6, diag, 1/3
      (3,4) GG: diag M(2,3)+5 = 5; up (3,3) U 1: -4; left (2,4) U 1: -9
                -> 5, diag, 1/4
      (4,2) AC: diag M(3,1)-4 = -4; up (4,1) out -> x; left (3,2) U 4: 0
                -> 0, left, 4/6
      (4,3) AA: diag M(3,2)+5 = 5; up (4,2) U 2: 0; left (3,3) U 1: -4
                -> 5, diag, 3/6
      (4,4) AG: diag M(3,3)-4 = 2; up (4,3) U 1: -5; left (3,4) U 1: -5
                -> 2, diag, 2/4
    Row 4: (4,2) 0, (4,3) 5, (4,4) 2 -> bestCol 5 (3/6).
    Column 4: (2,4) 1, (3,4) 5, (4,4) 2 -> bestRow 5 (1/4).
    5 > 5 false -> bestRow's cell: 1/4, score 5.  Charging GO + GE for the first
    gap makes (2,3)'s up 10 - 11 = -1 (M -1), so (3,4) = -1 + 5 = 4 and
    bestCol 5 > bestRow 4 takes bestCol's cell: 3/6."""
    case("H6", "ACAG", "ACGA", 2, 1, 4, 5)()


def test_raw_byte_compare():
    """s1 = "a", s2 = "A", band 2: kB = 2/4 = 0 (only the diagonal).  Bytes
    are compared as they are (:403, the read's own characters), so 'a' and
    'A' differ.

      (1,1) Aa: diag 0 - 4 = -4; up (1,0) out -> x; left (0,1) out -> x
                -> -4, diag, ML 0 - 0 + 1 = 1, AL 1 -> 1/1
    Row 1 and column 1: (1,1) -> both -4 (1/1); 1/1 = 1.0, score -4."""
    case("H7", "a", "A", 2, 1, 1, -4)()


def test_full_form_last_row_wins():
    """s1 = "ACG", s2 = "AT", band 1: the full form (genCPU.cpp:132-225):
    row 0 and column 0 are boundary cells in full (:147-161, :172-175) and
    every cell is computed.

      (1,1) AA: 5 / 0 / 0 -> 5, diag, 0/1
      (1,2) AC: diag M(0,1)-4 = -4; up (1,1) U 1: -5; left (0,2): 0
                -> 0, left, 0/2 + 1 = 1/3
      (1,3) AG: diag M(0,2)-4 = -4; up (1,2) U 2: 0; left (0,3): 0
                -> 0, up (before left), 1/3 + 1 = 2/4
      (2,1) TA: diag M(1,0)-4 = -4; up (2,0): 0; left (1,1) U 1: -5
                -> 0, up, 0/2 + 1 = 1/3
      (2,2) TC: diag M(1,1)-4 = 1; up (2,1) U 4: -1; left (1,2) U 2: -1
                -> 1, diag, 0 + 1 = 1 / 2 -> 1/2
      (2,3) TG: diag M(1,2)-4 = -4; up (2,2) U 1: -9; left (1,3) U 4: 0
                -> 0, left, 2/4 + 1 = 3/5
    Row 2: (2,1) 0, (2,2) 1, (2,3) 0 -> bestCol 1 (1/2).
    Column 3: (1,3) 0 -> bestRow 0 (2/4); (2,3) 0 >= 0 -> (3/5).
    1 > 0 -> bestCol's cell: 1/2, score 1."""
    case("F1", "ACG", "AT", 1, 1, 2, 1)()


def test_full_form_zigzag():
    """s1 = "AAAA", s2 = "CCC", band 1 (full form): no byte matches, so
    every step is a mismatch (-4) or a gap; gaps alternate direction for
    free after leaving the boundary.

      (1,1) CA: -4 / 0 / 0 -> 0, up, 1/2
      (1,2) CA: diag -4; up (1,1) U 4: -1; left (0,2): 0 -> 0, left, 0/2+1 = 1/3
      (1,3) CA: diag -4; up (1,2) U 2: 0; left (0,3): 0 -> 0, up, 2/4
      (1,4) CA: diag -4; up (1,3) U 4: -1; left (0,4): 0 -> 0, left, 1/5
      (2,1) CA: diag M(1,0)-4 = -4; up (2,0): 0; left (1,1) U 4: 0
                -> 0, up, 0/2 + 1 = 1/3
      (2,2) CA: diag M(1,1)-4 = -4; up (2,1) U 4: -1; left (1,2) U 2: -1
                -> -1, up, 2/4
      (2,3) CA: diag -4; up (2,2) U 4: -2; left (1,3) U 4: 0 -> 0, left, 3/5
      (2,4) CA: diag -4; up (2,3) U 2: 0; left (1,4) U 2: -1 -> 0, up, 4/6
      (3,1) CA: diag M(2,0)-4 = -4; up (3,0): 0; left (2,1) U 4: 0
                -> 0, up, 1/4
      (3,2) CA: diag M(2,1)-4 = -4; up (3,1) U 4: -1; left (2,2) U 4: -1
                -> -1, up, 2/5
      (3,3) CA: diag M(2,2)-4 = -5; up (3,2) U 4: -2; left (2,3) U 2: -1
                -> -1, left, 4/6
      (3,4) CA: diag M(2,3)-4 = -4; up (3,3) U 2: -1; left (2,4) U 4: 0
                -> 0, left, 5/7
    Row 3: (3,1) 0 (1/4), (3,2) -1, (3,3) -1, (3,4) 0 >= 0 -> bestCol 0 (5/7).
    Column 4: (1,4) 0 (1/5), (2,4) 0 (4/6), (3,4) 0 (5/7) -> bestRow 0 (5/7).
    0 > 0 false -> bestRow's cell: 5/7, score 0."""
    case("F2", "AAAA", "CCC", 1, 5, 7, 0)()


# ---------------------------------------------------------------------------
# Mutation check: each entry plants one plausible mistake in a copy of the
# oracle source (text substitutions that must each match exactly once) and
# names the hand case that must tell it apart.
# ---------------------------------------------------------------------------
_UP_PEN = "(U[cur][j - 1] & 1) * OR_GO + (U[cur][j - 1] >> 2 & 1) * OR_GE"
_LEFT_PEN = "(U[pre][j] & 1) * OR_GO + (U[pre][j] >> 1 & 1) * OR_GE"
MUTATIONS = {
    # the verdict's six
    "left after an up-cell charges GE": (
        "H1", [(_LEFT_PEN, "(U[pre][j] & 1) * OR_GO + ((U[pre][j] >> 1 | U[pre][j] >> 2) & 1) * OR_GE")]),
    "up after a left-cell charges GE": (
        "H2", [(_UP_PEN, "(U[cur][j - 1] & 1) * OR_GO + ((U[cur][j - 1] >> 2 | U[cur][j - 1] >> 1) & 1) * OR_GE")]),
    "final choice takes the row cell on a tie": (
        "H4", [("if (bestCol > bestRow) {", "if (bestCol >= bestRow) {")]),
    "column j = L1 keeps its first maximum": (
        "H1", [("if (j == L1 && m >= bestRow)", "if (j == L1 && m > bestRow)")]),
    "row i = L2 keeps its first maximum": (
        "H3", [("if (i == L2 && m >= bestCol)", "if (i == L2 && m > bestCol)")]),
    "up wins a diag/up tie": (
        "H5", [("if (m == diag) {", "if (m == diag && m != up) {")]),
    # further plausible slips
    "left wins an up/left tie": (
        "H1", [("} else if (m == up) {", "} else if (m == up && m != left) {")]),
    "left wins a diag/left tie": (
        "H8", [("if (m == diag) {", "if (m == diag && m != left) {")]),
    "up reads an out-of-band neighbour": (
        "H1", [("if (full || or_in_band(i, j - 1, kb))", "if (1)")]),
    "left reads an out-of-band neighbour": (
        "H3", [("if (full || or_in_band(i - 1, j, kb))", "if (1)")]),
    "boundary row AL = 0": (
        "H4", [("ML[cur][j] = 0; AL[cur][j] = j; }", "ML[cur][j] = 0; AL[cur][j] = 0; }")]),
    "boundary column AL = 0": (
        "H1", [("ML[cur][0] = 0; AL[cur][0] = i; }", "ML[cur][0] = 0; AL[cur][0] = 0; }")]),
    "boundary cells count as diagonals (gap from them pays GO)": (
        "H1", [("U[cur][j] = 0; ML[cur][j] = 0; AL[cur][j] = j; }", "U[cur][j] = 1; ML[cur][j] = 0; AL[cur][j] = j; }"),
               ("U[cur][0] = 0; ML[cur][0] = 0; AL[cur][0] = i; }", "U[cur][0] = 1; ML[cur][0] = 0; AL[cur][0] = i; }")]),
    "first gap pays GO + GE": (
        "H6", [(_UP_PEN, "(U[cur][j - 1] & 1) * (OR_GO + OR_GE) + (U[cur][j - 1] >> 2 & 1) * OR_GE"),
               (_LEFT_PEN, "(U[pre][j] & 1) * (OR_GO + OR_GE) + (U[pre][j] >> 1 & 1) * OR_GE")]),
    "diagonal ML counts matches as mismatches": (
        "H3", [("(ML[pre][j - 1] - isMatched) + 1", "ML[pre][j - 1] + 1")]),
    "reads swapped between rows and columns": (
        "H1", [("    int full = band <= 1;\n",
                "    { const char *t_ = s1; int l_ = L1; s1 = s2; L1 = L2; s2 = t_; L2 = l_; }\n"
                "    int full = band <= 1;\n")]),
    "kB = (L1 + L2) / band": (
        "H7", [("(L1 + L2) / (band * 2)", "(L1 + L2) / band")]),
    "band test is strict": (
        "H1", [("(i - j >= -kb) && (i - j <= kb)", "(i - j > -kb) && (i - j < kb)")]),
    "bytes compared case-insensitively": (
        "H7", [("int isMatched = s2[i - 1] == s1[j - 1];",
                "int isMatched = (s2[i - 1] | 32) == (s1[j - 1] | 32);")]),
    "full form initialises only kB boundary cells": (
        "F2", [("int row0 = full ? L1 : kb;", "int row0 = kb;")]),
}


def _build_mutant(name, subs, out_dir):
    src = open(ORACLE_SRC).read()
    for old, new in subs:
        assert src.count(old) == 1, f"{name}: pattern must occur once in the oracle: {old!r}"
        src = src.replace(old, new)
    tag = "".join(c if c.isalnum() else "_" for c in name)
    c_path = os.path.join(out_dir, tag + ".c")
    so_path = os.path.join(out_dir, tag + ".so")
    with open(c_path, "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-shared", "-w", "-o", so_path, c_path, "-lm"],
                   check=True)
    return so_path


def _case_result(lib, s1, s2, band):
    ml, al, sc = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    lib.oracle_nw_pair(s1, len(s1), s2, len(s2), band, ctypes.byref(ml), ctypes.byref(al),
                       ctypes.byref(sc))
    return ml.value, al.value, sc.value


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc builds the mutant oracles")
def test_hand_cases_kill_mutations(tmp_path):
    """Every planted mistake changes (ML, AL, score) of the case named for
    it, and the unmutated source passes all cases (sanity of the harness)."""
    for fn in (test_gap_leaving_boundary_then_free_switch, test_free_switch_up_after_left,
               test_last_row_keeps_last_maximum, test_final_tie_prefers_last_column,
               test_diag_wins_a_tie_with_up, test_diag_wins_a_tie_with_left,
               test_gap_open_after_diagonal_costs_go_only,
               test_raw_byte_compare, test_full_form_last_row_wins, test_full_form_zigzag):
        fn()  # fills CASES
    names = list(MUTATIONS)
    with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        libs = list(ex.map(lambda nm: _build_mutant(nm, MUTATIONS[nm][1], str(tmp_path)), names))
    clean = _build_mutant("unmutated", [], str(tmp_path))
    L = ctypes.CDLL(clean)
    L.oracle_nw_pair.restype = ctypes.c_float
    for cname, (s1, s2, band, ml, al, sc) in CASES.items():
        assert _case_result(L, s1, s2, band) == (ml, al, sc), cname
    survivors = []
    for nm, so in zip(names, libs):
        L = ctypes.CDLL(so)
        L.oracle_nw_pair.restype = ctypes.c_float
        killer = MUTATIONS[nm][0]
        s1, s2, band, ml, al, sc = CASES[killer]
        if _case_result(L, s1, s2, band) == (ml, al, sc):
            survivors.append((nm, killer))
    assert not survivors, f"mutations not told apart by their named case: {survivors}"


# The device DP may run transposed: rows over s1 and columns over s2 (the
# shorter read as rows costs fewer band rows).  Transposition turns the
# reference's "up" (same row) into the "left" (previous row) of the new frame
# and its last row into the new frame's last column.  The penalties, band
# and boundary are symmetric, so the reference's result comes out when the
# new frame (a) lets its "left" win a tie between the two gap moves and (b)
# takes its last row's best unless its last column's is strictly larger.
TRANSPOSED = [
    ("""            } else if (m == up) {
                U[cur][j] = 4; ML[cur][j] = ML[cur][j - 1] + 1; AL[cur][j] = AL[cur][j - 1] + 1;
            } else {
                U[cur][j] = 2; ML[cur][j] = ML[pre][j] + 1; AL[cur][j] = AL[pre][j] + 1;
            }""",
     """            } else if (m == left) {
                U[cur][j] = 2; ML[cur][j] = ML[pre][j] + 1; AL[cur][j] = AL[pre][j] + 1;
            } else {
                U[cur][j] = 4; ML[cur][j] = ML[cur][j - 1] + 1; AL[cur][j] = AL[cur][j - 1] + 1;
            }"""),
    ("    if (bestCol > bestRow) {", "    if (!(bestRow > bestCol)) {"),
]


def _tie_heavy_pairs(rng, count, max_len):
    for _ in range(count):
        alph = np.frombuffer(b"AGTC", np.uint8)[:int(rng.integers(1, 5))]
        l1, l2 = (int(x) for x in rng.integers(1, max_len + 1, 2))
        yield (alph[rng.integers(0, len(alph), l1)].tobytes(), alph[rng.integers(0, len(alph), l2)].tobytes(),
               int(rng.integers(1, 13)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc builds the transposed copies")
def test_transposed_frame_gives_the_reference_result(tmp_path):
    """A copy of the oracle with (a) and (b), called with the reads exchanged,
    returns the oracle's (ML, AL, score) on 6,000 random pairs over 1- to
    4-letter alphabets (ties everywhere) and lengths 1-40; each of (a) and (b)
    alone gives a different result on some of them."""
    built = {}
    for name, subs in (("transposed", TRANSPOSED), ("tie_only", TRANSPOSED[:1]), ("final_only", TRANSPOSED[1:])):
        L = ctypes.CDLL(_build_mutant(name, subs, str(tmp_path)))
        L.oracle_nw_pair.restype = ctypes.c_float
        built[name] = L
    ref = ctypes.CDLL(_build_mutant("reference_frame", [], str(tmp_path)))
    ref.oracle_nw_pair.restype = ctypes.c_float
    rng = np.random.default_rng(4242)
    differ = {"tie_only": 0, "final_only": 0}
    n = 0
    for s1, s2, band in _tie_heavy_pairs(rng, 6000, 40):
        want = _case_result(ref, s1, s2, band)
        assert _case_result(built["transposed"], s2, s1, band) == want, (s1, s2, band)
        for k in differ:
            differ[k] += _case_result(built[k], s2, s1, band) != want
        n += 1
    print("transposed frame: pairs where (a) or (b) alone differ", differ, "of", n)
    assert differ["tie_only"] > 0 and differ["final_only"] > 0, differ
