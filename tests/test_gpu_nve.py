"""NVE energy conservation of the mixed-precision path against the FP64 path
(the reference's arithmetic) from a common equilibrated start.

Moldy truncates the pair potentials at the cutoff without a shift, so H = KE
+ PE random-walks as pairs cross the cutoff whatever the precision; the bar
here is that mixed precision adds nothing visible to that: its rms deviation
about the linear fit and its largest excursion stay within 1.5x and 2x of the
FP64 run's over 1 ps.
"""
import math

import numpy as np
import pytest

from moldy_amd import system, units

pytestmark = pytest.mark.gpu


def _equilibrate(s, T=300.0, nblocks=8, nper=150, dt=0.001):
    from moldy_amd import engine as E
    e = E.ForceEngine(s, precision=E.FP64)
    e.set_dynamics(system.Dynamics(step=dt))
    e.upload_state()
    e.upload_momenta(*s.thermal_momenta(T, seed=7))
    kT = units.KB * T
    for _ in range(nblocks):
        tr, _, _ = e.md_steps(nper, init_H_0=True)
        d = e.download_dynamics()
        ft = math.sqrt(1.5 * s.nmols * kT / tr[-1, 0])
        fr = math.sqrt(1.5 * s.nmols_r * kT / tr[-1, 1])
        e.upload_momenta(d["mom"] * ft, d["amom"] * fr)
    d = e.download_dynamics()
    e.close()
    return d


def _nve(s, start, precision, nsteps=2000, dt=0.0005):
    from moldy_amd import engine as E
    e = E.ForceEngine(s, precision=precision)
    e.set_dynamics(system.Dynamics(step=dt))
    e.upload_state(c_of_m=start["c_of_m"], quat=start["quat"])
    e.upload_momenta(start["mom"], start["amom"])
    tr, _, _ = e.md_steps(nsteps, init_H_0=True)
    e.close()
    H = tr[:, :4].sum(1) * units.CONV_E / s.nmols
    t = np.arange(nsteps) * dt
    fit = np.polyfit(t, H, 1)
    return np.std(H - np.polyval(fit, t)), np.abs(H - H[0]).max(), fit[0] * 1000.0, tr


def test_mixed_nve_no_worse_than_fp64():
    from moldy_amd import engine as E
    s = system.water_box(512, seed=5, cutoff=9.0, alpha=0.30)
    start = _equilibrate(s)
    f64 = _nve(s, start, E.FP64)
    mix = _nve(s, start, E.MIXED)
    print(f"fp64: rms {f64[0]:.2e} max {f64[1]:.2e} drift {f64[2]:.2e} kJ/mol/mol/ns")
    print(f"mixed: rms {mix[0]:.2e} max {mix[1]:.2e} drift {mix[2]:.2e} kJ/mol/mol/ns")
    ke = f64[3][:, :2].sum(1).mean() * units.CONV_E / s.nmols
    assert f64[1] < 1e-2 * ke          # the FP64 run itself conserves H to 1% of KE
    assert mix[0] < 1.5 * f64[0]
    assert mix[1] < 2.0 * f64[1]
