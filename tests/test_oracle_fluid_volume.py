"""Pins of the constant-fluid-volume oracle (oracle/fluid_volume.py) on the CPU.

The reference's own driver for this method (test_DippingCrackConstantVolume.jl) asserts nothing, and the
minimiser it calls (Optim's Brent) is not vendored, so the pins are mathematical:
* Brent's method and the golden-section search find the minimisers of functions with known minima to their
  tolerance; Brent is exact on a parabola after its first interpolation step.
* A penny crack under fluid pressure alone must hold the target volume: |V - volume| / V below the bound that
  Brent's abs_tol = 0.001 on the pressure factor x in [4, 10] allows, every crack element open, no shear slip.
* The same with a normal compression that keeps part of [0, 10] closed.
* More than one fracture ends in the reference's error.
"""
import numpy as np
import pytest

from oracle import fluid_volume as FV
from oracle import oracle as O
from tests import stl

G_, NU_ = 1.0, 0.25


def penny():
    P, T = stl.read_stl("VertPenny-300-EqEdges.stl")
    P1, P2, P3 = (P[T[:, k] - 1, 1:4].copy() for k in range(3))
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    return P1, P2, P3, nrm, mid


def normal_stress(nrm, s):
    """Remote stress with normal traction -s on every element of a planar crack (normal along the mean normal)."""
    n = nrm.shape[0]
    m = nrm.mean(axis=0)
    m /= np.linalg.norm(m)
    sig = -s * np.outer(m, m)
    return tuple(np.full(n, sig[i, j]) for i, j in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)))


def test_brent_and_golden_section_on_known_minima():
    x, fx, calls = FV.brent(lambda t: (t - 2.0) ** 2 + 1.0, 0.0, 10.0, abs_tol=0.001)
    assert abs(x - 2.0) < 1e-9 and fx == pytest.approx(1.0) and calls <= 8
    x, _, _ = FV.brent(lambda t: np.exp(t) - 3 * t, 0.0, 10.0, abs_tol=1e-6)
    assert abs(x - np.log(3.0)) < 1e-5
    x, _, _ = FV.brent(lambda t: t, 1.0, 5.0, abs_tol=1e-4)  # minimum at the lower bound
    assert x - 1.0 < 1e-3
    x, _, calls = FV.golden_section(lambda t: (t - 7.5) ** 2, 0.0, 50.0)
    assert abs(x - 7.5) < 1e-6 and calls > 30


def _solve(s):
    P1, P2, P3, nrm, mid = penny()
    n = P1.shape[0]
    lam = 2 * G_ * NU_ / (1 - 2 * NU_)
    V = np.pi * 1.0 ** 2 * 0.01
    z = np.zeros(n)
    out = FV.slip_fluid_volume(P1, P2, P3, NU_, G_, lam, mid, nrm, 0, normal_stress(nrm, s), (z, z, z), np.ones(n), V)
    return out, V


@pytest.mark.parametrize("s", [0.0, 0.004])
def test_penny_holds_target_volume(s):
    (Dn, Dss, Dds, info), V = _solve(s)
    assert info["method"] == "brent"
    assert 0.0 < info["x"] <= 10.0
    # |x - x*| <= ~2 (rel_tol |x| + abs_tol) with the volume affine in x for an open crack: volume error relative
    # to the pressure-driven part at most 2.5 (1e-3 + 1.5e-8 x) / x*, x* >= 10 / 2.5 once Norm is settled
    assert abs(info["volume"] - V) / V < 2e-3
    assert np.all(Dn > 0)
    assert np.max(np.abs(Dss)) < 1e-6 * np.max(Dn) and np.max(np.abs(Dds)) < 1e-6 * np.max(Dn)


def test_more_than_one_fracture_is_the_reference_error():
    P1, P2, P3, nrm, mid = penny()
    n = P1.shape[0]
    flag = np.ones(n)
    flag[: n // 2] = 2
    z = np.zeros(n)
    with pytest.raises(FV.VolumeError, match="Check scl here"):
        FV.slip_fluid_volume(P1, P2, P3, NU_, G_, 0.4, mid, nrm, 0, (z,) * 6, (z, z, z), flag, 1.0)
