"""The Moldy-signature drop-ins (libmoldyforce.so: force_calc, ewald, poteval)
called with Moldy's own data layout, against the oracle."""
import numpy as np
import pytest

from conftest import rel_err
from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["water", "clay", "silica", "argon_strict", "water_molpbc"])
def test_shim_force_calc_and_ewald(case):
    from moldy_amd.moldy_abi import MoldyCall
    from oracle.oracle import Oracle
    s = CASES[case]()
    o = Oracle(s)
    xyz = o.make_sites()
    MoldyCall.release()
    mc = MoldyCall(s)
    mc.set_sites(xyz)
    f, pe, st = mc.force_calc()
    ref = o.force_calc(xyz)
    assert rel_err(f, ref["force"]) < 1e-10
    assert abs(pe - ref["pe"]) <= 1e-10 * abs(ref["pe"])
    assert rel_err(np.triu(st), np.triu(ref["stress"])) < 1e-10
    if s.control.alpha > 0:
        f, pe, st = mc.ewald()
        ref = o.ewald(xyz)
        assert rel_err(f, ref["force"]) < 1e-10
        assert abs(pe - ref["pe"]) <= 1e-10 * abs(ref["pe"])
        assert rel_err(np.triu(st), np.triu(ref["stress"])) < 1e-10
    MoldyCall.release()
