"""Pins of the oracle's do_step (src/accel.c:626-827, src/leapfrog.c) by what
the mathematics fixes, on the CPU:

* free flight: with no interactions the centre of mass moves by
  step * h^-1 p / m per step, exactly;
* symmetric top: for two equal moments of inertia the splitting of
  leapf_quat_b is exact (src/leapfrog.c:249-254), so n steps of dt equal one
  step of n*dt;
* asymmetric free rotor: agreement with a fine RK4 integration of Euler's
  equations and dq/dt = q (0, w) / 2, with error falling as dt^2 -- this fixes
  the direction conventions of the rotation sub-steps;
* Nose-Poincare: (H - H_0) s is conserved to O(dt^2) (Bond, Leimkuhler and
  Laird 1999), with H = KE + PE + p_s^2 / 2Q + g kT ln s;
* barostat: with the cell terms |p_h|^2 / 2W + P det(h), the same extended
  Hamiltonian (or plain H at constant temperature off) is conserved to
  O(dt^2) for Parrinello-Rahman and uniform dilation, over an ideal gas,
  free rotors and a cold crystal whose virial drives the cell; uniform
  dilation keeps the shape of h and the strain mask freezes its entries.
"""
import math

import numpy as np
import pytest

from moldy_amd import system, units
from oracle.oracle import Oracle


def free_molecules(p_f, masses, nmols=8, L=30.0, seed=4, rdof=3):
    """nmols copies of one molecule with every interaction switched off."""
    rng = np.random.Generator(np.random.Philox(seed))
    nsite = len(p_f)
    ids = list(range(1, nsite + 1))
    max_id = nsite + 1
    mass = np.zeros(max_id)
    mass[1:] = masses
    sp = system.Species("free", nmols, ids, np.asarray(p_f, np.float64), False, rdof)
    ctl = system.Control(cutoff=6.0, alpha=-1.0)
    g = int(math.ceil(nmols ** (1 / 3)))
    grid = np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3)[:nmols]
    pos = (grid + 0.5) / g - 0.5
    q = system.random_quaternions(rng, nmols) if rdof else np.zeros((0, 4))
    return system.MoldySystem([sp], mass, np.zeros(max_id), system.empty_potpar(max_id), 0, ctl,
                              np.eye(3) * L, pos, q, "free")


def water_rotor(n=4):
    return free_molecules(system.tip4p_sites(), [system.M_O, system.M_H, system.M_H, 0.0], n)


def symmetric_top(n=4):
    # three equal masses on an equilateral triangle: I_x = I_y = I_z / 2
    r = 1.1
    p = np.array([[r * math.cos(a), r * math.sin(a), 0.0] for a in (0, 2 * math.pi / 3, 4 * math.pi / 3)])
    return free_molecules(p, [12.0, 12.0, 12.0], n)


def run(s, dyn, nsteps, T=300.0, seed=2):
    o = Oracle(s)
    mom, amom = s.thermal_momenta(T, seed=seed)
    step, st = o.md(dyn, mom, amom)
    recs = []
    for i in range(nsteps):
        rc, pe, stress, rec = step(i == 0)
        assert rc == 0, rc
        recs.append(rec.copy())
    return st, np.array(recs), (mom, amom)


def test_free_flight():
    s = free_molecules([[0.0, 0.0, 0.0]], [20.0], nmols=27, rdof=0)
    dt, n = 0.002, 7
    st, _, (mom, _) = run(s, system.Dynamics(step=dt), n)
    hinv = np.linalg.inv(s.h)
    v = mom @ hinv.T @ hinv.T / 20.0        # h^-1 p / m with p = h^-T mom
    expect = s.c_of_m + n * dt * v
    expect -= np.floor(expect + 0.5)
    assert np.abs(st["c_of_m"] - expect).max() < 1e-12
    assert np.abs(st["mom"] - mom).max() == 0.0


def test_symmetric_top_splitting_is_exact():
    s = symmetric_top()
    mass, inertia = s.species_mass_inertia()
    assert abs(inertia[0, 0] - inertia[0, 1]) < 1e-12 * inertia[0, 0]
    a, _, _ = run(s, system.Dynamics(step=0.004), 4)
    b, _, _ = run(s, system.Dynamics(step=0.016), 1)
    assert np.abs(a["quat"] - b["quat"]).max() < 1e-12
    assert np.abs(a["amom"] - b["amom"]).max() < 1e-10 * np.abs(b["amom"]).max()


def _euler_rk4(q, L, inertia, t, nsub):
    """Free rigid rotor: body-frame dL/dt = L x w, dq/dt = q (0, w) / 2."""
    def qmul(p, r):
        p0, p1, p2, p3 = p
        r0, r1, r2, r3 = r
        return np.array([p0 * r0 - p1 * r1 - p2 * r2 - p3 * r3, p1 * r0 + p0 * r1 - p3 * r2 + p2 * r3,
                         p2 * r0 + p3 * r1 + p0 * r2 - p1 * r3, p3 * r0 - p2 * r1 + p1 * r2 + p0 * r3])

    def f(y):
        q, L = y[:4], y[4:]
        w = L / inertia
        return np.r_[0.5 * qmul(q, np.r_[0.0, w]), np.cross(L, w)]
    y = np.r_[q, L]
    h = t / nsub
    for _ in range(nsub):
        k1 = f(y); k2 = f(y + 0.5 * h * k1); k3 = f(y + 0.5 * h * k2); k4 = f(y + h * k3)
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return y[:4] / np.linalg.norm(y[:4]), y[4:]


@pytest.mark.parametrize("nosym", [False, True])
def test_asymmetric_rotor_matches_euler_equations(nosym):
    s = water_rotor(3)
    _, inertia = s.species_mass_inertia()
    o = Oracle(s)
    mom, amom = s.thermal_momenta(300.0, seed=5)
    T = 0.02
    errs = []
    for nstep in (20, 40):
        step, st = o.md(system.Dynamics(step=T / nstep, nosymmetric_rot=nosym), mom, amom)
        for i in range(nstep):
            assert step(i == 0)[0] == 0
        err = 0.0
        for m in range(3):
            q_ref, L_ref = _euler_rk4(o.f["quat"].reshape(-1, 4)[m], amom[m, 1:], inertia[0], T, 4000)
            err = max(err, np.abs(st["quat"][m] - q_ref).max(),
                      np.abs(st["amom"][m, 1:] - L_ref).max() / np.abs(L_ref).max())
        errs.append(err)
    assert errs[0] < 1e-3
    # second order: halving dt divides the error by about 4
    assert 2.5 < errs[0] / errs[1] < 6.0, errs


def test_free_rotor_energy_second_order():
    s = water_rotor(8)
    spans = []
    for dt in (0.002, 0.001):
        _, rec, _ = run(s, system.Dynamics(step=dt), int(0.04 / dt))
        e = rec[:, 0] + rec[:, 1]
        spans.append(np.ptp(e) / e.mean())
    assert spans[0] < 1e-5
    assert 2.5 < spans[0] / spans[1] < 6.0, spans


@pytest.mark.parametrize("kind", ["particles", "rotors"])
def test_nose_poincare_invariant(kind):
    if kind == "particles":
        s, Q = free_molecules([[0.0, 0.0, 0.0]], [20.0], nmols=27, rdof=0), 5.0
    else:
        s, Q = water_rotor(8), 50.0
    T = 300.0
    gkt = s.degrees_of_freedom() * units.KB * T
    spans = []
    for dt in (0.002, 0.001):
        dyn = system.Dynamics(step=dt, temp=T, const_temp=1, ttmass=Q)
        _, rec, _ = run(s, dyn, int(0.04 / dt), T=400.0)
        ke, pe, ts, tsmom, H0 = rec[:, 0] + rec[:, 1], rec[:, 2] + rec[:, 3], rec[:, 4], rec[:, 5], rec[:, 6]
        # the record point has half of the potential kick on p but none of it
        # on p_s (accel.c:713-760): add that half to place p_s consistently
        ps = tsmom - 0.5 * dt * (pe - H0)
        H = ke + pe + ps ** 2 / (2 * Q) + gkt * np.log(ts)
        spans.append(np.ptp((H - H0) * ts) / ke.mean())
        assert np.ptp(ts) > 1e-3          # the thermostat does act
    assert spans[0] < 1e-3
    assert 2.5 < spans[0] / spans[1] < 6.0, spans


# ---- barostat (do_step's const_pressure branches: gleap_cell, update_hmom,
# stress_kin; src/accel.c:686-703, 761-779, src/leapfrog.c:415-531) ----

def _hmom_kick(h, stress, s, dt, uniform):
    """The cell-momentum increment of update_hmom for a stress times volume."""
    if uniform:
        return dt * s * np.trace(stress) / (h ** 2).sum() * h
    return stress @ np.linalg.inv(h).T * (dt * s)


def _barostat_run(s, dyn, tt, T):
    """The extended Hamiltonian of the record point of each step.

    At the record point p has half of the potential kick but p_s and the cell
    momentum have none of theirs (accel.c:713-761): add those halves.  The
    reference's stress carries the impulsive term of the truncated potential,
    (distp - dist) / V on the diagonal (force.c distant-pair correction); it is
    the volume derivative of (distp - dist) / V, which therefore joins the
    energy."""
    o = Oracle(s)
    mom, amom = s.thermal_momenta(T, seed=3)
    step, st = o.md(dyn, mom, amom)
    cx = o.distant_const(1) - o.distant_const(0)
    gkt = s.degrees_of_freedom() * units.KB * dyn.temp
    mask = np.array([(dyn.strain_mask >> k) & 1 for k in range(9)]).reshape(3, 3) == 1
    H, vol, hs = [], [], []
    dt = dyn.step
    for i in range(int(round(tt / dt))):
        rc, pe, stress, rec = step(i == 0)
        assert rc == 0
        ke, pot, ts, tsmom, H0, V = rec[0] + rec[1], rec[2] + rec[3], rec[4], rec[5], rec[6], rec[7]
        hm, h = rec[8:17].reshape(3, 3), rec[17:26].reshape(3, 3)
        assert abs(np.linalg.det(h) - V) < 1e-9 * V
        hmc = hm + 0.5 * _hmom_kick(h, stress, ts, dt, dyn.const_pressure == 2)
        hmc[mask] = 0.0
        Hx = ke + pot + (hmc ** 2).sum() / (2 * dyn.pmass) + dyn.pressure * V + cx / V
        if dyn.const_temp:
            ps = tsmom - 0.5 * dt * (pot - H0)
            Hx = (Hx + ps ** 2 / (2 * dyn.ttmass) + gkt * np.log(ts) - H0) * ts
        H.append(Hx)
        vol.append(V)
        hs.append(h)
    return np.array(H), np.array(vol), np.array(hs), ke, st


def _barostat_cases():
    P_gas = 4.0 / units.CONV_P            # about N k T / V of the gas below
    return {
        "gas_nph": (lambda: free_molecules([[0.0, 0.0, 0.0]], [20.0], nmols=27, rdof=0),
                    dict(temp=300.0, pmass=0.5, pressure=P_gas)),
        "gas_npt": (lambda: free_molecules([[0.0, 0.0, 0.0]], [20.0], nmols=27, rdof=0),
                    dict(temp=300.0, pmass=0.5, pressure=P_gas, const_temp=1, ttmass=5.0)),
        "rotors_npt": (lambda: water_rotor(8),
                       dict(temp=300.0, pmass=0.5, pressure=P_gas, const_temp=1, ttmass=50.0)),
        # cold fcc argon with the cutoff between neighbour shells, so that no
        # pair crosses it: the virial drives the cell
        "crystal_nph": (lambda: system.argon_fcc(ncell=4, jitter=0.02, cutoff=8.6),
                        dict(temp=10.0, pmass=500.0, pressure=50.0 / units.CONV_P)),
    }


@pytest.mark.parametrize("const_pressure", [1, 2])
@pytest.mark.parametrize("case", list(_barostat_cases()))
def test_barostat_invariant_second_order(case, const_pressure):
    build, kw = _barostat_cases()[case]
    spans = []
    for dt in (0.004, 0.002) if case == "crystal_nph" else (0.002, 0.001):
        dyn = system.Dynamics(step=dt, const_pressure=const_pressure, **kw)
        H, vol, hs, ke, _ = _barostat_run(build(), dyn, 0.04, kw["temp"])
        spans.append(np.ptp(H) / ke)
        assert np.ptp(vol) / vol.mean() > 2e-4      # the cell does move
    assert spans[0] < 5e-5, spans
    assert 3.0 < spans[0] / spans[1] < 5.0, spans


def test_barostat_cell_constraints():
    """Uniform dilation keeps the shape of h; the strain mask freezes its
    entries (default 200: the lower triangle); mask 511 freezes the cell."""
    build, kw = _barostat_cases()["rotors_npt"]
    s = build()
    s.h = s.h @ np.array([[1.0, 0.05, 0.02], [0.0, 1.0, 0.03], [0.0, 0.0, 1.0]])
    h0 = s.h.copy()
    _, _, hs, _, st = _barostat_run(s, system.Dynamics(step=0.002, const_pressure=2, **kw), 0.02,
                                    kw["temp"])
    nz = h0 != 0
    ratio = hs[-1][nz] / h0[nz]
    assert np.ptp(ratio) < 1e-12 and abs(ratio[0] - 1.0) > 1e-4
    assert np.all(hs[-1][~nz] == 0.0)
    _, _, hs, _, _ = _barostat_run(s, system.Dynamics(step=0.002, const_pressure=1, **kw), 0.02,
                                   kw["temp"])
    assert np.all(hs[:, [1, 2, 2], [0, 0, 1]] == 0.0)
    assert np.abs(hs[-1] - h0)[0, 1] > 1e-6           # upper triangle is free
    mask = 200 | 1 | 16 | 256                         # diagonal frozen too
    _, _, hs, _, _ = _barostat_run(s, system.Dynamics(step=0.002, const_pressure=1, strain_mask=mask,
                                                      **kw), 0.02, kw["temp"])
    assert np.all(hs[:, [0, 1, 2], [0, 1, 2]] == h0[[0, 1, 2], [0, 1, 2]])
    # all frozen, no thermostat: the trajectory is the constant-volume one
    # (with the thermostat the P V terms of gleap_cell shift p_s at the H_0
    # point, so H_0 and the trajectory differ)
    kw = dict(kw, const_temp=0)
    a = _barostat_run(s, system.Dynamics(step=0.002, const_pressure=1, strain_mask=511, **kw), 0.02,
                      kw["temp"])[4]
    b = _barostat_run(s, system.Dynamics(step=0.002, **kw), 0.02, kw["temp"])[4]
    for k in ("c_of_m", "quat", "mom", "amom"):
        assert np.allclose(a[k], b[k], rtol=0, atol=1e-12), k
