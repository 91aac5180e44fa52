"""Host helpers of the drop-in (no GPU).

* TD_sum (TD_sum.jl:1-41) applied to influence matrices reproduces the summed-field form of TD (TD.jl:219-410),
  as the reference's Okada test uses it (test/test_TDvsOkada.jl:100): the oracle's influence matrices for a
  two-triangle fault with slip go in, the oracle's summed fields are the check.
* An empty strain or displacement set passes through unchanged.
"""
import numpy as np

from cutanddisplacejulia_amd import api
from oracle import analytic as A
from oracle import oracle as O


def _fields():
    P1, P2, P3, _ = A.okada_test_fault()
    xx = np.linspace(-6, 6, 9)
    X, Y = np.meshgrid(xx, xx)
    x, y = X.ravel(), Y.ravel()
    z = -0.3 - 0.1 * np.abs(x)
    Dss, Dds, Dn = [0.5, -0.2], [2.0, 1.1], [0.6, 0.3]
    return x, y, z, P1, P2, P3, Dss, Dds, Dn


def test_td_sum_equals_summed_field():
    x, y, z, P1, P2, P3, Dss, Dds, Dn = _fields()
    s, d = O.td(x, y, z, P1, P2, P3, Dss, Dds, Dn, 0.25, 1.0, 1, 1, 1)
    out = api.TD_sum(*[s[k] for k in O.STRAIN_FIELDS], *[d[k] for k in O.DISP_FIELDS])
    sf, df = O.td_field(x, y, z, P1, P2, P3, Dss, Dds, Dn, 0.25, 1.0, 1, 1, 1)
    for got, ref in zip(out, list(sf) + list(df)):
        assert got.shape == (x.size, 1)
        assert np.max(np.abs(got[:, 0] - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_td_sum_passes_empty_sets_through():
    x, y, z, P1, P2, P3, Dss, Dds, Dn = _fields()
    s, _ = O.td(x, y, z, P1, P2, P3, Dss, Dds, Dn, 0.25, 1.0, 0, 1, 0)
    e = []
    out = api.TD_sum(*[s[k] for k in O.STRAIN_FIELDS], *([e] * 9))
    assert all(o is e for o in out[6:])
    assert all(o.shape == (x.size, 1) for o in out[:6])
    out = api.TD_sum(*([e] * 18), *([np.ones((3, 2))] * 9))
    assert all(o is e for o in out[:6])
    assert all(np.array_equal(o, np.full((3, 1), 6.0)) for o in out[6:])
