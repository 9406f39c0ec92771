"""The row groups' cross-lane scan, replayed on the CPU against its recurrence.

In a row group (osqp_amd/csrc/serve/team.h, group_rows) the part of a row
carried in from earlier lanes follows C_l = b_l + a_l C_{l-1} (a_l in {0, 1}),
and the kernel solves it with a fixed sequence of DPP moves (scan_step: row
shifts within rows of 16 lanes, then row broadcasts under row masks).  This
test reads that sequence and the DPP control values from team.h itself, models
each move as __builtin_amdgcn_update_dpp does with bound_ctrl off (a lane whose
source is out of range, or whose row the mask leaves out, keeps the `old`
operand), and checks the result equals the sequential recurrence on random
and structured lane patterns.
"""
import os
import random
import re

from .conftest import ROOT

TEAM_H = os.path.join(ROOT, "osqp_amd", "csrc", "serve", "team.h")


def _schedule():
    src = open(TEAM_H).read()
    consts = {k: int(v, 16) for k, v in re.findall(r"(kDpp\w+) = (0x[0-9a-fA-F]+)", src)}
    body = src[src.index("void group_rows("):]
    body = body[:body.index("\n}\n")]
    steps = re.findall(r"scan_step<(kDpp\w+)(?:,\s*(0x[0-9a-fA-F]+))?>\(a, b\);", body)
    return [(consts[c], int(m, 16) if m else 0xF) for c, m in steps]


def _dpp(old, v, ctrl, row_mask):
    out = []
    for lane in range(64):
        row = lane // 16
        src = None
        if (row_mask >> row) & 1:
            if 0x111 <= ctrl <= 0x11F:          # row_shr:n
                d = ctrl - 0x110
                src = lane - d if lane % 16 >= d else None
            elif ctrl == 0x142:                 # row_bcast:15 (lane 15 of the row before)
                src = row * 16 - 1 if row >= 1 else None
            elif ctrl == 0x143:                 # row_bcast:31 (lane 31 to rows 2 and 3)
                src = 31 if row >= 2 else None
            else:
                raise AssertionError(f"DPP control {ctrl:#x} not modelled")
        out.append(v[src] if src is not None else old)
    return out


def _scan(a, b, schedule):
    A, Bv = list(a), list(b)
    for ctrl, mask in schedule:
        ap, bp = _dpp(1, A, ctrl, mask), _dpp(0, Bv, ctrl, mask)
        Bv = [Bv[i] + bp[i] if A[i] else Bv[i] for i in range(64)]
        A = [A[i] & ap[i] for i in range(64)]
    return Bv


def _recurrence(a, b):
    out, c = [], 0
    for lane in range(64):
        c = b[lane] + (a[lane] * c if lane else 0)
        out.append(c)
    return out


def test_schedule_read_from_source():
    sched = _schedule()
    # four row shifts and two row broadcasts: the six steps of a 64-lane scan
    assert len(sched) == 6 and [c for c, _ in sched[:4]] == [0x111, 0x112, 0x114, 0x118], sched


def test_scan_matches_recurrence():
    sched = _schedule()
    rng = random.Random(2024)
    patterns = [[0] + [1] * 63, [0] * 64, [0] + [1, 0] * 31 + [1]]
    patterns += [[0] + [int(rng.random() < p) for _ in range(63)] for p in (0.3, 0.6, 0.9) for _ in range(1000)]
    for a in patterns:
        b = [rng.randint(-9, 9) for _ in range(64)]
        assert _scan(a, b, sched) == _recurrence(a, b), (a, b)
    # a dropped step must be caught
    assert any(_scan(a, b, sched[:-1]) != _recurrence(a, b)
               for a in patterns[:50] for b in [[rng.randint(-9, 9) for _ in range(64)]])
