"""Duplicate removal keyed by interaction types (typedDedup = 1, an extension), in the oracle, CPU.
The reference keeps the first path (launch ascending, breadth-first order) per receiver and
coplanar-id sequence (RTSolver.cpp:435-459), so a path that reflects where another transmits at the
same surfaces is dropped.  Keyed by ids and types, the groups split: the paths kept are a superset
of the reference's, every extra path shares its id sequence with a path the reference keeps but not
its type sequence, per receiver every (ids, types) key is kept once, in the reference's order, and
the receiver's field is the reference's plus the extra paths' fields."""
import numpy as np

from oracle import oracle as O
from tests.parity import oracle_paths, path_table
from tests.office import KW, MATS
from tests.office import scene_and_rx as _rx


def _per_rx(table):
    out = {}
    for rxi, launch, tri, cop, typ in table:
        out.setdefault(rxi, []).append((launch, cop, typ))
    return out


def test_typed_groups_split_the_reference_groups():
    sc, rx = _rx()
    for kw in (dict(KW), dict(KW, tess=4, maxT=2)):
        a = O.trace(sc, O.make_params(**kw), rx, MATS, nthreads=8)
        b = O.trace(sc, O.make_params(**dict(kw, typedDedup=1)), rx, MATS, nthreads=8)
        ta, tb = path_table(oracle_paths(a)), path_table(oracle_paths(b))
        assert set(ta) <= set(tb) and len(tb) > len(ta)
        ra, rb = _per_rx(ta), _per_rx(tb)
        for r, paths in rb.items():
            keys = [(cop, typ) for _, cop, typ in paths]
            assert len(set(keys)) == len(keys)                      # one path per (ids, types)
            assert [p[0] for p in paths] == sorted(p[0] for p in paths)   # launch order kept
            ref = {cop: typ for _, cop, typ in ra.get(r, [])}
            assert len(ref) == len(ra.get(r, []))                   # reference: one per ids
            for _, cop, typ in paths:
                assert cop in ref                                   # same id groups as the reference
        # a receiver's field is the reference's plus the fields of its extra paths
        sa = set(ta)
        extra = np.zeros_like(b.fields)
        for i, row in enumerate(tb):
            if row not in sa:
                extra[row[0]] += b.path_fields[i]
        assert np.abs(extra).max() > 0
        scale = np.abs(a.fields).max(axis=1) + np.abs(extra).max(axis=1)
        assert np.all(np.abs(b.fields - (a.fields + extra)).max(axis=1) <= 1e-9 * scale)
