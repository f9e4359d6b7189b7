"""The oracle's renormalize (nuosc_ana.cpp:140-166) pinned by the identities
its form fixes, not by restating it.  Per point, with
P~ = (emr, -emi, ee - mm) / G0 (P1 without the factor 2 that analysis()
uses) and P~b = (bemr, bemi, bee - bmm) / G0b, the new values are

    ee' + mm'  = 2 G0                    bee' + bmm'  = 2 G0b
    ee' - mm'  = 2 G0 P~3 / |P~|         bee' - bmm'  = 2 G0b P~b3 / |P~b|
    emr' = emr / |P~|,  emi' = emi / |P~|   (and the bar pair with |P~b|)

so the trace doubles and the polarisation keeps its direction.  A
renormalisation with the factor 2 restored in P1 (the form analysis() uses)
is shown to fail the same check."""
import numpy as np
import pytest


def random_state(o, rng):
    nx, nz, nv = o.shape
    st = np.empty((8, nx, nz, nv))
    G0, G0b = o.G0, o.G0b
    for base, g in ((0, G0), (4, G0b)):
        p = rng.uniform(-1.0, 1.0, (3, nx, nz, nv)) * rng.uniform(0.2, 1.3, (1, nx, nz, nv))
        st[base + 0] = 0.5 * g * (1.0 + p[2])
        st[base + 1] = 0.5 * g * (1.0 - p[2])
        st[base + 2] = 0.5 * g * p[0]
        st[base + 3] = 0.5 * g * p[1]
    return st


def identity_errors(before, after, G0, G0b):
    """Max relative deviation from each identity (G0 broadcast over v)."""
    errs = {}
    for base, g, s2 in ((0, G0, -1.0), (4, G0b, 1.0)):
        ee, mm, emr, emi = before[base:base + 4]
        ee2, mm2, emr2, emi2 = after[base:base + 4]
        P = np.sqrt(emr ** 2 + (s2 * emi) ** 2 + (ee - mm) ** 2) / g
        tag = "nu" if base == 0 else "nubar"
        errs[tag + " trace"] = np.max(np.abs(ee2 + mm2 - 2 * g)) / np.max(2 * g)
        errs[tag + " P3"] = np.max(np.abs((ee2 - mm2) - 2 * (ee - mm) / P)) / np.max(2 * g)
        errs[tag + " emr"] = np.max(np.abs(emr2 - emr / P)) / np.max(np.abs(emr / P))
        errs[tag + " emi"] = np.max(np.abs(emi2 - emi / P)) / np.max(np.abs(emi / P))
    return errs


def renorm_factor2(st, G0, G0b):
    """A plausible mistake, on the CPU: P1 = 2 emr / G0 (analysis()'s form)."""
    out = st.copy()
    for base, g in ((0, G0), (4, G0b)):
        ee, mm, emr, emi = st[base:base + 4]
        P = np.sqrt((2 * emr) ** 2 + emi ** 2 + (ee - mm) ** 2) / g
        t = (ee - mm) / g / P * g
        out[base], out[base + 1] = g + t, g - t
        out[base + 2], out[base + 3] = emr / P, emi / P
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_renormalize_identities(oracle_mod, seed):
    O = oracle_mod
    o = O.OracleSim(nn=3, x0=-0.3, x1=0.3, z0=-0.4, z1=0.4, dx=0.1, dz=0.1, renorm=1)
    o.fill_init(ipt=3, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5)
    G0, G0b = o.G0, o.G0b
    st = random_state(o, np.random.default_rng(seed))
    o.set_state(st)
    o.renormalize()
    after = o.get_state()
    errs = identity_errors(st, after, G0, G0b)
    assert max(errs.values()) <= 4e-15, errs
    # the mutated form fails the check by far
    bad = identity_errors(st, renorm_factor2(st, G0, G0b), G0, G0b)
    assert bad["nu emr"] > 1e-2 and bad["nu P3"] > 1e-2, bad
    o.close()


def test_renormalize_twice_keeps_trace_and_phase(oracle_mod):
    """After one call P~' = (P~1, P~2, 2 P~3) / |P~|, whose norm is not 1 in
    general, so a second call changes the state again: the reference's form
    is not a projection.  What every call keeps: the trace at 2 G0 and the
    ratio emr/emi."""
    O = oracle_mod
    o = O.OracleSim(nn=2, x0=-0.2, x1=0.2, z0=-0.3, z1=0.3, dx=0.1, dz=0.1, renorm=1)
    o.fill_init(ipt=3, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5)
    G0 = o.G0
    st = random_state(o, np.random.default_rng(5))
    o.set_state(st)
    o.renormalize()
    o.renormalize()
    after = o.get_state()
    assert np.max(np.abs(after[0] + after[1] - 2 * G0)) <= 4e-15 * np.max(G0)
    ratio0 = st[2] / st[3]
    ratio2 = after[2] / after[3]
    assert np.max(np.abs(ratio2 - ratio0) / np.abs(ratio0)) <= 1e-13
    o.close()
