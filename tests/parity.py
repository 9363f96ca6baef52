"""Parity measures between two unbinding results (GPU vs oracle, or two GPU modes)."""
import numpy as np


def bound_sets(out, pop):
    """Per subhalo: array of local indices of bound particles."""
    res = out["result"]
    sets = []
    for s in range(pop.nsub):
        a = pop.offsets[s]
        nb = int(res[s]["nbound"])
        sets.append(out["order"][a:a + nb])
    return sets


def bound_mass_agreement(a, b, pop):
    """Mass of the symmetric difference of the bound sets over the reference bound mass.

    Returns (summed fraction, per-subhalo fractions); b is the reference."""
    sa, sb = bound_sets(a, pop), bound_sets(b, pop)
    tot_diff = 0.0
    tot_ref = 0.0
    per = np.zeros(pop.nsub)
    for s in range(pop.nsub):
        base = pop.offsets[s]
        m = pop.mass[base:pop.offsets[s + 1]].astype(np.float64)
        x = np.zeros(len(m), bool)
        y = np.zeros(len(m), bool)
        x[sa[s]] = True
        y[sb[s]] = True
        d = m[x ^ y].sum()
        r = m[y].sum()
        tot_diff += d
        tot_ref += r
        per[s] = d / r if r > 0 else (0.0 if d == 0 else np.inf)
    return (tot_diff / tot_ref if tot_ref > 0 else 0.0), per


def energy_errors(a, b, pop, pot_ref):
    """|E_a - E_b| / |potential_b| over particles bound in both (b is the reference;
    pot_ref = reference ParticlePotentialEnergies in b's final order)."""
    errs = []
    for s in range(pop.nsub):
        base = pop.offsets[s]
        na = int(a["result"][s]["nbound"])
        nb = int(b["result"][s]["nbound"])
        if na == 0 or nb == 0:
            continue
        ia = a["order"][base:base + na]
        ib = b["order"][base:base + nb]
        ea = np.full(pop.offsets[s + 1] - base, np.nan)
        ea[ia] = a["energy"][base:base + na]
        eb = np.full_like(ea, np.nan)
        eb[ib] = b["energy"][base:base + nb]
        pb = np.full_like(ea, np.nan)
        pb[ib] = pot_ref[base:base + nb]
        both = ~np.isnan(ea) & ~np.isnan(eb)
        errs.append(np.abs(ea[both] - eb[both]) / np.abs(pb[both]))
    return np.concatenate(errs) if errs else np.zeros(0)


def order_is_reference_layout(out, pop, max_sample_size=0, refine_fraction=0.1):
    """Bound part ascending in energy (after the centre-refined prefix, whose
    stored energies are kept out of order, src/subhalo_unbind.cpp:730-733), and
    every particle present once."""
    for s in range(pop.nsub):
        a, b = pop.offsets[s], pop.offsets[s + 1]
        nb = int(out["result"][s]["nbound"])
        e = out["energy"][a:b]
        if out["result"][s]["disrupted"]:
            continue
        start = 0
        if max_sample_size > 0 and nb > max_sample_size:
            start = min(nb, max(max_sample_size, int(np.float32(refine_fraction) * np.float32(nb))))
        if nb - start > 1 and np.any(np.diff(e[start:nb]) < 0):
            return False
        if sorted(out["order"][a:b].tolist()) != list(range(b - a)):
            return False
    return True


def order_differences_are_ties(a, b, pop, rtol=2e-5):
    """Compare the final orders of a and b slot by slot.  Where they differ, the
    two particles must be energy ties in b (within rtol of the subhalo's largest
    |E|): the reference sorts with the unstable std::sort
    (src/subhalo_unbind.cpp:673,710), so the order of ties is unpinned, and the
    GPU's FP32 energies can swap particles closer than its rounding.  Disrupted
    subhaloes keep their input order on both sides.  Returns the number of slots
    that differ; raises AssertionError on any other difference."""
    ndiff = 0
    for s in range(pop.nsub):
        lo, hi = pop.offsets[s], pop.offsets[s + 1]
        if hi == lo:
            continue  # an orphan: no particles on either side
        oa, ob = a["order"][lo:hi], b["order"][lo:hi]
        da, db = bool(a["result"][s]["disrupted"]), bool(b["result"][s]["disrupted"])
        assert da == db, f"subhalo {s}: disrupted {da} vs {db}"
        if da:
            assert np.array_equal(oa, ob), f"subhalo {s}: disrupted order differs"
            continue
        eb = np.empty(hi - lo, np.float64)
        eb[ob] = b["energy"][lo:hi]
        scale = rtol * max(np.abs(eb).max(), 1e-30)
        k = np.where(oa != ob)[0]
        ndiff += len(k)
        bad = k[np.abs(eb[oa[k]] - eb[ob[k]]) > scale]
        assert len(bad) == 0, (f"subhalo {s}: slots {bad[:5]} hold {oa[bad[:5]]} vs {ob[bad[:5]]} with energies "
                               f"{eb[oa[bad[:5]]]} vs {eb[ob[bad[:5]]]}")
    return ndiff
