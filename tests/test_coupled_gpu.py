"""GPU checks of the self-consistent thermo-electrical loop (meta.shockley
ThermoElectric3D on a seeded layered device): power balance (Joule heat = I U,
the identity the reference's shockley2d.py:60-61 test asserts), heating above the
heat-sink temperature, and FP32-operator + FP64-refinement PCG agreeing with FP64
PCG on the final temperatures."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(shape, prec, outer):
    from plask_amd import coupled
    te, info = coupled.build(shape=shape)
    te.thermal.iterative.precision = prec
    te.electrical.iterative.precision = prec
    r = coupled.run(te, outer)
    return te, info, r


def test_thermoelectric_power_balance_and_precision_agreement():
    shape = (24, 24, 48)
    print("coupled fp64", shape, flush=True)
    te, info, r64 = _run(shape, 0, 6)
    I = te.get_total_current()
    P = te.electrical.get_total_heat()
    print("I = %.6g mA, U = %.4f V, P = %.6g mW, I U = %.6g mW, Tmax %.4f" % (I, info["V"], P, abs(I) * info["V"],
                                                                          r64["Tmax"]), flush=True)
    # current flows from the top contact down (j = -sigma grad V < 0 along z)
    assert abs(I) > 0
    assert abs(P - abs(I) * info["V"]) <= 0.01 * abs(I) * info["V"]
    assert r64["Tmax"] > 300.0
    T64 = te.thermal._temps().copy()
    print("coupled fp32 refine", shape, flush=True)
    te32, _, r32 = _run(shape, 1, 6)
    T32 = te32.thermal._temps()
    d = np.abs(T32 - T64).max() / np.abs(T64).max()
    print("max rel T difference %.3e" % d, flush=True)
    assert d <= 1e-6
