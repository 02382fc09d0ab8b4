"""RDF bins on the GPU (mgpu_rdf) against the oracle's rdf_inner/rdf_accum
restatement: bit-identical float bins, including accumulation into bins that
already hold values (Moldy accumulates over RDF steps)."""
import numpy as np
import pytest

from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", list(CASES))
def test_rdf_bitwise(case):
    from moldy_amd import engine as E
    from oracle.oracle import Oracle
    s = CASES[case]()
    o = Oracle(s)
    xyz = o.make_sites()
    limit = min(0.45 * float(np.diag(s.h).min()), 10.0)
    nbins = 100
    rc, ref = o.rdf(xyz, limit, nbins)
    assert rc == 0
    e = E.ForceEngine(s)
    got = e.rdf(s.h, xyz, limit, nbins)
    print(case, "pairs binned:", int(np.rint(ref.sum() * s.nsites / np.linalg.det(s.h))))
    assert np.array_equal(got, ref)
    # second RDF step accumulates
    rc, ref2 = o.rdf(xyz, limit, nbins, rdf=ref.copy())
    got2 = e.rdf(s.h, xyz, limit, nbins, rdf=got)
    assert np.array_equal(got2, ref2)
    e.close()
