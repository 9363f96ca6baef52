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
