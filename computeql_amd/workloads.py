"""Synthetic workloads for the batched and served forms.

event_table(rows, seed): the 30-column event table (host arrays, seeded):
    0  user_id   u32  Zipf(s=1.1) over 1M users
    1  status    u32  uniform over 5
    2  region    u32  uniform over 50
    3  category  u32  dictionary codes of 200 distinct strings
    4-11  m0..m7 f64  lognormal(mu=3, sigma=1)
    12-21 t0..t9 i64  timestamps (t0-t4) and counters below 2^32 (t5-t9)
    22-29 u0..u7 u32  uniform over [0, 1e6)
event_nulls(rows, seed, frac) flags NULL rows of the nullable columns, the
timestamps t0-t4 (frac of each, 2% by default); the mixed queries read them
only in their WHERE, where a NULL row fails its comparison.

mixed_queries(n, seed, rows): n filter / GROUP BY aggregates over that table
with literals drawn from the seed; shapes cycle through the list below.
"""
from __future__ import annotations

import numpy as np

from .engine import AVG, COUNT, F64, I64, MAX, MIN, SUM, U32, AggSpec, Table
from .ir import col

USER, STATUS, REGION, CATEGORY = 0, 1, 2, 3
M0, T0, U0 = 4, 12, 22
EVENT_COLUMNS = ([("user_id", U32), ("status", U32), ("region", U32), ("category", U32)]
                 + [(f"m{i}", F64) for i in range(8)] + [(f"t{i}", I64) for i in range(10)]
                 + [(f"u{i}", U32) for i in range(8)])
CATEGORY_NAMES = [f"cat-{i:03d}" for i in range(200)]
TS_BASE = 1_700_000_000_000


def event_table(rows: int, seed: int = 7) -> list:
    rng = np.random.default_rng(seed)
    cols = [((rng.zipf(1.1, rows) - 1) % 1_000_000).astype(np.uint32),
            rng.integers(0, 5, rows).astype(np.uint32),
            rng.integers(0, 50, rows).astype(np.uint32),
            rng.integers(0, 200, rows).astype(np.uint32)]
    cols += [rng.lognormal(3.0, 1.0, rows) for _ in range(8)]
    cols += [TS_BASE + rng.integers(0, 86_400_000 * 30, rows, dtype=np.int64) for _ in range(5)]
    # counters below 2^32: a SUM over any of the mixed queries' groups stays
    # below 2^53, where the engine fails the query (query_exec.c:7028-7049)
    cols += [rng.integers(0, 1 << 32, rows, dtype=np.int64) for _ in range(5)]
    cols += [rng.integers(0, 1_000_000, rows).astype(np.uint32) for _ in range(8)]
    return cols


NULLABLE = tuple(range(T0, T0 + 5))


def event_nulls(rows: int, seed: int = 7, frac: float = 0.02) -> dict:
    """{column: byte-per-row flags, nonzero = NULL} for the nullable columns."""
    rng = np.random.default_rng(seed ^ 0x5EED)
    return {c: (rng.random(rows) < frac).astype(np.uint8) for c in NULLABLE}


def upload(cols: list, nulls: dict | None = None) -> Table:
    t = Table(len(cols[0]))
    for i, ((name, ty), c) in enumerate(zip(EVENT_COLUMNS, cols)):
        t.add_column(name, ty, c, nulls=None if nulls is None else nulls.get(i))
    return t


def mixed_queries(n: int, seed: int = 11) -> list:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        shape = i % 8
        m = M0 + int(rng.integers(0, 8))
        t = T0 + int(rng.integers(0, 5))
        c = T0 + 5 + int(rng.integers(0, 5))
        u = U0 + int(rng.integers(0, 8))
        if shape == 0:
            out.append(AggSpec([COUNT], [None], (col(STATUS) == int(rng.integers(0, 5)))
                               & (col(REGION) < int(rng.integers(1, 50)))))
        elif shape == 1:
            out.append(AggSpec([COUNT, SUM, AVG], [None, m, m], col(u) < int(rng.integers(1, 1_000_000)), STATUS))
        elif shape == 2:
            lo = TS_BASE + int(rng.integers(0, 86_400_000 * 20))
            out.append(AggSpec([COUNT, SUM], [None, c], (col(t) >= lo) & (col(t) < lo + 86_400_000 * 3), REGION))
        elif shape == 3:
            out.append(AggSpec([MIN, MAX, COUNT], [m, m, None], col(CATEGORY) == int(rng.integers(0, 200))))
        elif shape == 4:
            out.append(AggSpec([COUNT], [None], col(REGION) == int(rng.integers(0, 50)), CATEGORY))
        elif shape == 5:
            out.append(AggSpec([SUM, MAX], [u, u], col(m) > float(rng.lognormal(3.0, 1.0)), STATUS))
        elif shape == 6:
            out.append(AggSpec([COUNT, SUM, MIN, MAX], [None, m, m, m],
                               (col(STATUS) < int(rng.integers(1, 5))) & (col(u) >= int(rng.integers(0, 500_000)))))
        else:
            out.append(AggSpec([COUNT, AVG], [None, u], (col(REGION) >= int(rng.integers(0, 40)))
                               & (col(CATEGORY) < int(rng.integers(1, 200))), REGION))
    return out
