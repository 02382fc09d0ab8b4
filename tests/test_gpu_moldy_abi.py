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


@pytest.mark.parametrize("case", ["water", "clay"])
def test_shim_rdf_on_rdf_steps(case):
    """force_calc accumulates the RDF into Moldy's rdf_ptr() array on RDF steps
    (src/force.c:1302-1313) and leaves it alone on the others."""
    import ctypes as C
    import os
    from conftest import RDF_STUB
    from moldy_amd.moldy_abi import MoldyCall, ContrMt, shim
    from oracle.oracle import Oracle
    if not os.path.exists(RDF_STUB):
        pytest.skip("rdf_ptr stand-in not built (make -C tests/abi_stub)")
    stub = C.CDLL(RDF_STUB, mode=C.RTLD_GLOBAL)
    s = CASES[case]()
    o = Oracle(s)
    xyz = o.make_sites()
    limit, nbins = min(0.45 * float(np.diag(s.h).min()), 9.0), 60
    nblk = (s.max_id - 1) * s.max_id // 2
    buf = np.zeros((nblk, nbins), np.float32)
    stub.stub_set_rdf(buf.ctypes.data_as(C.POINTER(C.c_float)), buf.size)
    MoldyCall.release()
    mc = MoldyCall(s)
    ctl = ContrMt.in_dll(shim(), "control")
    ctl.rdf_interval, ctl.begin_rdf, ctl.limit, ctl.nbins = 5, 10, limit, nbins
    mc.set_sites(xyz)
    for istep in (5, 10, 11, 15):
        ctl.istep = istep
        mc.force_calc()
    _, ref = o.rdf(xyz, limit, nbins)
    _, ref = o.rdf(xyz, limit, nbins, rdf=ref)        # steps 10 and 15
    ctl.rdf_interval = 0
    stub.stub_set_rdf(None, 0)
    MoldyCall.release()
    assert np.array_equal(buf, ref)
