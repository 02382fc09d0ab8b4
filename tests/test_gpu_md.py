"""The device MD step (do_step, src/accel.c:626-827) against the oracle's
restatement, and graph replay against eager launches.

FP64 trajectories agree with the oracle to 1e-9 relative over a few steps
(the same arithmetic up to summation order, amplified by a short chaotic
horizon); graph-replayed steps are bit-identical to eagerly launched ones.
"""
import numpy as np
import pytest

from conftest import rel_err
import systems_small as S
from moldy_amd import system

pytestmark = pytest.mark.gpu

TOL = 1e-9


def md_cases():
    return {
        "water_nve": (lambda: S.water(), system.Dynamics(step=0.0005)),
        "water_nvt": (lambda: S.water(), system.Dynamics(step=0.0005, temp=300.0, const_temp=1,
                                                          ttmass=50.0)),
        "water_nve_quat_a": (lambda: S.water(), system.Dynamics(step=0.0005, nosymmetric_rot=True)),
        "water_nvt_quat_a": (lambda: S.water(), system.Dynamics(step=0.0005, temp=300.0, const_temp=1,
                                                                 ttmass=50.0, nosymmetric_rot=True)),
        "salt_nve": (lambda: S.salt(), system.Dynamics(step=0.0005)),
        "argon_nvt": (lambda: S.argon(), system.Dynamics(step=0.005, temp=120.0, const_temp=1,
                                                          ttmass=20.0)),
        "water_molpbc": (lambda: _molpbc(S.water()), system.Dynamics(step=0.0005)),
    }


def _molpbc(s):
    s.control.molpbc = True
    return s


def _engine(s, dyn, mom, amom, precision=0):
    from moldy_amd import engine as E
    e = E.ForceEngine(s, precision=precision)
    e.set_dynamics(dyn)
    e.upload_state()
    e.upload_momenta(mom, amom)
    return e


@pytest.mark.parametrize("case", list(md_cases()))
@pytest.mark.parametrize("graph", [False, True])
def test_md_steps_vs_oracle(case, graph):
    from oracle.oracle import Oracle
    build, dyn = md_cases()[case]
    s = build()
    mom, amom = s.thermal_momenta(350.0 if "argon" not in case else 120.0, seed=3)
    nsteps = 6
    o = Oracle(s)
    step, st = o.md(dyn, mom, amom)
    ref_rec = []
    for i in range(nsteps):
        rc, pe, stress, rec = step(i == 0)
        assert rc == 0
        ref_rec.append(rec[:7])
    ref_rec = np.array(ref_rec)
    e = _engine(s, dyn, mom, amom)
    tr, pe_g, st_g = e.md_steps(nsteps, init_H_0=True, use_graph=graph)
    got = e.download_dynamics()
    errs = dict(c_of_m=rel_err(got["c_of_m"], st["c_of_m"]), mom=rel_err(got["mom"], st["mom"]),
                trace=rel_err(tr[:, :7], ref_rec), pe=rel_err(pe_g, pe), stress=rel_err(st_g, stress))
    if s.nmols_r:
        errs["quat"] = rel_err(got["quat"], st["quat"])
        errs["amom"] = rel_err(got["amom"], st["amom"])
    errs["ts"] = abs(got["ts"] - st["ts"])
    errs["tsmom"] = abs(got["tsmom"] - st["tsmom"]) / max(abs(st["tsmom"]), 1.0)
    print(case, "graph" if graph else "eager", {k: f"{v:.1e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < TOL, (k, v)
    if dyn.const_temp:
        assert abs(st["ts"] - 1.0) > 1e-8          # the thermostat moved s
    e.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_is_bitwise_eager(precision):
    s = S.water()
    dyn = system.Dynamics(step=0.0005, temp=300.0, const_temp=1, ttmass=50.0)
    mom, amom = s.thermal_momenta(300.0, seed=4)
    out = []
    for graph in (False, True):
        e = _engine(s, dyn, mom, amom, precision)
        tr1, _, _ = e.md_steps(5, init_H_0=True, use_graph=graph)
        tr2, _, _ = e.md_steps(7, init_H_0=False, use_graph=graph)   # replays an existing graph
        out.append((np.vstack([tr1, tr2]), e.download_dynamics()))
        e.close()
    (ta, da), (tb, db) = out
    assert np.array_equal(ta, tb)
    for k in ("c_of_m", "quat", "mom", "amom"):
        assert np.array_equal(da[k], db[k]), k


def test_graph_rebuilt_after_state_upload():
    """A new upload_state with another h invalidates the captured graph."""
    s = S.water()
    dyn = system.Dynamics(step=0.0005)
    mom, amom = s.thermal_momenta(300.0, seed=4)
    e = _engine(s, dyn, mom, amom)
    e.md_steps(3, init_H_0=True)
    h2 = s.h * 1.01
    e.upload_state(h=h2)
    e.upload_momenta(mom, amom)
    tr_g, _, _ = e.md_steps(4, init_H_0=True, use_graph=True)
    e.upload_state(h=h2)
    e.upload_momenta(mom, amom)
    tr_e, _, _ = e.md_steps(4, init_H_0=True, use_graph=False)
    assert np.array_equal(tr_g, tr_e)
    e.close()
