"""Comparison helpers shared by the parity tests (HIP library against the FP64 oracle)."""
import numpy as np

FIELD_RTOL = 1e-4   # per-receiver complex field, relative to |E| of the oracle (0.01 dB class)
# Absolute floor, relative to the strongest entry: a field that is zero in exact arithmetic (a
# transmission through the reference's infinite wall keeps only the radial component, about
# 1e-17 of the incident field) is rounding noise on both sides and has no meaningful relative error.
FIELD_ATOL_REL = 1e-12


def path_table(p):
    """List of (rx, launch, tri sequence, coplanar sequence, type sequence) per kept path."""
    out = []
    for k in range(len(p["prx"])):
        o, n = int(p["poff"][k]), int(p["pnseg"][k])
        out.append((int(p["prx"][k]), int(p["plaunch"][k]), tuple(p["stri"][o:o + n].tolist()),
                    tuple(p["scop"][o:o + n].tolist()), tuple(p["stype"][o:o + n].tolist())))
    return out


def oracle_paths(r):
    return dict(prx=r.prx, plaunch=r.plaunch, pnseg=r.pnseg, poff=r.poff, stri=r.stri, scop=r.scop, stype=r.stype,
                ssrc=r.ssrc, stgt=r.stgt)


def compare(orc, gpu_res, label=""):
    """Asserts identical kept-path sets (as ordered triangle-id sequences, per receiver, in the
    reference's order), bit-identical corrected vertices, equal bounce counts and fields within
    FIELD_RTOL.  Returns the worst relative field error."""
    gp = gpu_res.paths()
    op = oracle_paths(orc)
    assert gpu_res.bounces == orc.bounces, f"{label}: bounces {gpu_res.bounces} != {orc.bounces}"
    tg, to = path_table(gp), path_table(op)
    if tg != to:
        sg, so = set(tg), set(to)
        missing = sorted(so - sg)[:5]
        extra = sorted(sg - so)[:5]
        raise AssertionError(f"{label}: path sets differ: {len(tg)} gpu vs {len(to)} oracle; "
                             f"missing {missing} extra {extra}")
    assert np.array_equal(gp["ssrc"], op["ssrc"]) and np.array_equal(gp["stgt"], op["stgt"]), \
        f"{label}: corrected vertices differ (max {np.abs(gp['ssrc'] - op['ssrc']).max()})"
    pf_g, pf_o = gpu_res.path_fields(), orc.path_fields
    worst_path = worst_field_error(pf_g, pf_o, f"{label} (per path)", to)
    worst = worst_field_error(gpu_res.fields, orc.fields, f"{label} (per receiver)")
    return max(worst, worst_path)


def worst_field_error(eg, eo, label, table=None):
    """Asserts per-receiver (or per-path) fields within FIELD_RTOL of the oracle's, with NaN on the
    same entries; returns the worst relative error among entries above -180 dB of the strongest."""
    # NaN where the reference produces NaN (the half-wave pattern on its axis is 0/0): the HIP
    # path must produce NaN on exactly the same entries.
    ng, no = ~np.isfinite(eg).all(axis=1), ~np.isfinite(eo).all(axis=1)
    assert np.array_equal(ng, no), f"{label}: non-finite entries differ: gpu {np.flatnonzero(ng)[:5]} " \
                                    f"oracle {np.flatnonzero(no)[:5]}"
    eg, eo = eg[~no], eo[~no]
    mag = np.linalg.norm(eo, axis=1)
    err = np.linalg.norm(eg - eo, axis=1)
    if mag.size == 0:
        return 0.0
    atol = FIELD_ATOL_REL * mag.max()
    bad = err > FIELD_RTOL * mag + atol
    if np.any(bad):
        i = int(np.flatnonzero(bad)[np.argmax((err / np.maximum(mag, 1e-300))[bad])])
        extra = f" path {table[np.flatnonzero(~no)[i]]}" if table is not None else ""
        raise AssertionError(f"{label}: field err {err[i]} > {FIELD_RTOL} * {mag[i]} + {atol} at {i}: "
                             f"gpu {eg[i]} oracle {eo[i]}" + extra)
    sig = mag > atol / FIELD_ATOL_REL * 1e-9     # entries above -180 dB of the strongest
    return float((err[sig] / mag[sig]).max()) if np.any(sig) else 0.0
