"""The FP32 form (FP32 storage and stencil/interaction arithmetic, FP64
compensated moments and diagnostics) against the FP64 path, and both against
the closed-form linear growth rate.

Stated tolerances:
  * FP64 growth rate of |rho_ex| = eigenvalue to 1e-5 (relative);
  * FP32 growth rate within 1% of the FP64 one;
  * FP32 Pee(t) within 1e-6 of FP64 and FP32 max|P|-1 below 1e-5 over the run.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import growth as G  # noqa: E402

pytestmark = pytest.mark.gpu

COL = dict(nv=4, x0=0.0, x1=0.1, z0=-3.2, z1=3.2, dx=0.1, dz=0.1, CFL=0.4, ko=0.1, mu=1.0, pmo=0.0)


def run_mode(N, precision, st, nblocks=12, per=25):
    g = N.NuOsc(precision=precision, **COL)
    g.fillInitValue(ipt=3, alpha=0.9, eps0=0.0, sigma=1.0, lnue=0.6, lnueb=0.5)
    g.set_state(st)
    ts, rs, pee, dp = [], [], [], []
    for b in range(nblocks):
        a = g.analysis()
        ts.append(a.phy_time)
        rs.append(a.rho_ex2)
        pee.append(a.surv)
        dp.append(a.maxdP)
        g.step_rk4(per)
    return g, np.array(ts), np.array(rs), np.array(pee), np.array(dp)


def test_growth_rate_fp64_and_fp32(nuosc_mod):
    N = nuosc_mod
    g = N.NuOsc(**COL)
    g.fillInitValue(ipt=3, alpha=0.9, eps0=0.0, sigma=1.0, lnue=0.6, lnueb=0.5)
    G0, G0b = g.g0()
    vg = g.vgrid()
    Z = g.coords()[1]
    st, gamma = G.mode_eigenstate(vg, G0, G0b, 1.0, 3, Z, 0.1, 0.1, 1e-6)
    _, t64, r64, p64, d64 = run_mode(N, N.PREC_FP64, st)
    _, t32, r32, p32, d32 = run_mode(N, N.PREC_FP32, st)
    g64 = G.fitted_rate(t64, r64)
    g32 = G.fitted_rate(t32, r32)
    print(f"growth rate: eigen {gamma:.10f} fp64 {g64:.10f} fp32 {g32:.10f} "
          f"(fp32 rel err {abs(g32 - g64) / g64:.2e}); Pee max err {np.max(np.abs(p32 - p64)):.2e}; "
          f"fp32 max|P|-1 {np.max(d32):.2e}")
    assert abs(g64 - gamma) <= 1e-5 * gamma
    assert abs(g32 - g64) <= 1e-2 * g64
    assert np.max(np.abs(p32 - p64)) <= 1e-6
    assert np.max(d32) <= 1e-5
