"""Pins of the oracle's RDF (rdf_inner, src/force.c:1010-1103; rdf_accum,
src/rdf.c:94-108) by brute force: every site pair closer than the limit,
found by an O(N^2) minimum-image scan (limit < half the box), counted once,
framework-framework pairs excluded; the float bins equal the counts added
one 1/density at a time in float."""
import numpy as np
import pytest

import systems_small as S
from oracle.oracle import Oracle


def brute_counts(s, xyz, limit, nbins):
    ids = s.site_id_array()
    fw = np.concatenate([np.full(sp.nmols * sp.nsites, sp.framework) for sp in s.species])
    hinv = np.linalg.inv(s.h)
    nblk = (s.max_id - 1) * s.max_id // 2
    counts = np.zeros((nblk, nbins), np.int64)
    blk = {}
    k = 0
    for a in range(1, s.max_id):
        for b in range(a, s.max_id):
            blk[(a, b)] = blk[(b, a)] = k
            k += 1
    r = xyz.T
    for i in range(s.nsites - 1):
        d = r[i + 1:] - r[i]
        f = d @ hinv.T
        d = (f - np.round(f)) @ s.h.T
        rr = np.sqrt((d * d).sum(1))
        js = np.nonzero((rr < limit) & ~(fw[i] & fw[i + 1:]))[0]
        for j, x in zip(js + i + 1, rr[js]):
            b = int(nbins / limit * x)
            if b < nbins:
                counts[blk[(ids[i], ids[j])], b] += 1
    return counts


def float_sum(counts, invrho):
    out = np.zeros(counts.shape, np.float32)
    flat = out.reshape(-1)
    for q, n in enumerate(counts.reshape(-1)):
        x = np.float32(0.0)
        for _ in range(n):
            x = np.float32(np.float64(x) + invrho)
        flat[q] = x
    return out


@pytest.mark.parametrize("case", ["water", "argon", "silica"])
def test_rdf_matches_brute_force(case):
    s = {"water": lambda: S.water(n=64, cutoff=5.0), "argon": lambda: S.argon(),
         "silica": lambda: S.silica()}[case]()
    limit = min(0.45 * np.diag(s.h).min(), 7.0)
    nbins = 40
    o = Oracle(s)
    xyz = o.make_sites()
    rc, rdf = o.rdf(xyz, limit, nbins)
    assert rc == 0
    counts = brute_counts(s, xyz, limit, nbins)
    invrho = 1.0 / (s.nsites / np.linalg.det(s.h))
    assert counts.sum() > 1000
    assert np.array_equal(rdf, float_sum(counts, invrho))


def test_rdf_accumulates():
    s = S.water(n=64, cutoff=5.0)
    o = Oracle(s)
    xyz = o.make_sites()
    _, one = o.rdf(xyz, 6.0, 30)
    _, two = o.rdf(xyz, 6.0, 30, rdf=one.copy())
    counts = np.rint(one * (s.nsites / np.linalg.det(s.h))).astype(np.int64)
    assert np.array_equal(two, float_sum(2 * counts, 1.0 / (s.nsites / np.linalg.det(s.h))))
