"""nuosc_p3_slice_noalloc (no device allocation; the slice staged in the
library's fixed per-GPU scratch array) against nuosc_p3_slice, bit for bit,
on two calls in a row with the second slice smaller, in FP64, FP32 and with
3 flavours; slices beyond the array's limits are refused before any GPU
work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(nv=3, x0=-0.8, x1=0.8, z0=-1.6, z1=1.6, dx=0.1, dz=0.1, CFL=0.4, ko=0.1, mu=1.0, pmo=1.0)


@pytest.mark.parametrize("precision,nflavor", [(64, 2), (32, 2), (64, 3)])
def test_noalloc_slice_equals_p3_slice(nuosc_mod, precision, nflavor):
    N = nuosc_mod
    g = N.NuOsc(precision=precision, nflavor=nflavor, **CFG)
    g.fillInitValue(ipt=5, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=77)
    g.step_rk4(4)
    big = (list(range(g.nx)), list(range(g.nv)))
    small = ([0, g.nx // 2, g.nx - 1], [3, 4, 5])
    for xs, vs in (big, small):
        want = g.p3_slice(xs, vs)
        got = g.p3_slice(xs, vs, noalloc=True)
        assert got.shape == (len(xs), g.nz, len(vs))
        assert np.array_equal(got, want)
        assert np.all(np.isfinite(got)) and np.max(np.abs(got)) > 0.5
    # the state moves on: the next call sees the new state
    g.step_rk4(2)
    assert np.array_equal(g.p3_slice(*small, noalloc=True), g.p3_slice(*small))
    g.close()


def test_noalloc_slice_limits(nuosc_mod):
    N = nuosc_mod
    g = N.NuOsc(**CFG)
    g.fillInitValue(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)
    # repeated x indices: 2000 * 32 * 42 values exceed the 2088960-value output region
    with pytest.raises(N.NuOscError) as e:
        g.p3_slice([0] * 2000, list(range(g.nv)), noalloc=True)
    assert e.value.code == N.NUOSC_EINVAL
    # more than 16384 indices
    with pytest.raises(N.NuOscError) as e:
        g.p3_slice([0] * 17000, [0], noalloc=True)
    assert e.value.code == N.NUOSC_EINVAL
    with pytest.raises(N.NuOscError):
        g.p3_slice([g.nx], [0], noalloc=True)
    # an empty slice is fine, and a slice at the limit still works
    assert g.p3_slice([], [0], noalloc=True).shape == (0, g.nz, 1)
    xs = [1] * (2088960 // (g.nz * g.nv))
    assert np.array_equal(g.p3_slice(xs, list(range(g.nv)), noalloc=True), g.p3_slice(xs, list(range(g.nv))))
    g.close()
