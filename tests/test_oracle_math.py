"""Pins the oracle on what the mathematics fixes, independently of its code:
the A&S erf against math.erf, every potential form against its definition
(energies) and numerical derivatives (forces), the link-cell sum against an
O(N^2) minimum-image sum, Newton's third law, the NaCl Madelung constant,
reciprocal forces and virial against finite differences, cellbin's edge
clamp, and the neighbour-cell offsets against a brute-force enumeration.
"""
import math

import numpy as np
import pytest

from moldy_amd import system, units
import systems_small as S


def oracle(s):
    from oracle.oracle import Oracle
    return Oracle(s)


def test_err_fn_is_abramowitz_stegun():
    from oracle.oracle import lib
    L = lib()
    for x in np.linspace(0, 5, 201):
        assert abs(L.orc_err_fn(x) - math.erf(x)) <= 1.5e-7   # A&S 7.1.26 bound
        assert abs(L.orc_err_fn(-x) + math.erf(x)) <= 1.5e-7


# potential definitions (src/kernel.c:46-59 comments, in Moldy's parameter order)
DEFS = {
    0: lambda r, p: p[0] * ((p[1] / r) ** 12 - (p[1] / r) ** 6),
    1: lambda r, p: -p[0] / r ** 6 + p[1] * np.exp(-p[2] * r),
    2: lambda r, p: p[0] * np.exp(-p[1] * r) - p[2] * np.exp(-p[3] * r),
    3: lambda r, p: p[0] * np.exp(-p[1] * r) + p[2] / r ** 12 - p[3] / r ** 4 - p[4] / r ** 6
    - p[5] / r ** 8,
    4: lambda r, p: p[0] / r ** 4 + p[1] / r ** 6 + p[2] / r ** 12,
    6: lambda r, p: p[0] * np.exp((p[1] - r) * p[2]) - p[3] / r ** 6
    + p[4] * (np.exp(-2 * p[5] * (r - p[6])) - 2 * np.exp(-p[5] * (r - p[6]))),
}
PARS = {0: [2.5, 3.1], 1: [150.0, 5e4, 3.4], 2: [6e4, 3.4, 150.0, 1.7],
        3: [6e4, 3.4, 8e4, 40.0, 250.0, 300.0], 4: [40.0, 250.0, 8e4],
        6: [2.5e4, 0.4, 3.1, 150.0, 4.0, 1.4, 3.1]}


def _pair(ptype, r, qq, alpha, p):
    import ctypes as C
    from oracle.oracle import lib
    pp = np.zeros(8)
    pp[:len(p)] = p
    rsq = np.array([r * r])
    nq = np.array([qq])
    f = np.zeros(1)
    pe = C.c_double(0)
    cols = (C.POINTER(C.c_double) * 8)(*[
        (pp[k:k + 1].ctypes.data_as(C.POINTER(C.c_double))) for k in range(8)])
    norm = 2 * alpha / math.sqrt(math.pi)
    rc = lib().orc_kernel(0, 1, f.ctypes.data_as(C.POINTER(C.c_double)), C.byref(pe),
                          rsq.ctypes.data_as(C.POINTER(C.c_double)),
                          nq.ctypes.data_as(C.POINTER(C.c_double)), C.c_double(1.0),
                          C.c_double(norm), C.c_double(alpha), ptype, cols)
    assert rc == 0
    return pe.value, f[0]


@pytest.mark.parametrize("ptype", sorted(DEFS))
@pytest.mark.parametrize("alpha", [-1.0, 0.35])
def test_pair_forms_energy_and_force(ptype, alpha):
    p = PARS[ptype]
    qq = 37.0 if alpha > 0 else 0.0
    for r in [2.2, 2.9, 3.7, 5.0, 7.5]:
        u, f = _pair(ptype, r, qq, alpha, p)
        coul = qq * math.erfc(alpha * r) / r if alpha > 0 else 0.0
        # energy: the definition, with the A&S erfc (|error| <= 1.5e-7)
        assert abs(u - (DEFS[ptype](r, p) + coul)) <= 1e-10 * abs(DEFS[ptype](r, p)) + 1.6e-7 * abs(qq) / r
        # force: -U'(r)/r of the non-Coulomb part by central differences; the
        # Coulomb part is qq (erfc(ar)/r + 2a/sqrt(pi) exp(-a^2 r^2)) / r^2
        hstep = 1e-5
        dU = (DEFS[ptype](r + hstep, p) - DEFS[ptype](r - hstep, p)) / (2 * hstep)
        fc = qq * (math.erfc(alpha * r) / r + 2 * alpha / math.sqrt(math.pi)
                   * math.exp(-(alpha * r) ** 2)) / r ** 2 if alpha > 0 else 0.0
        ref = -dU / r + fc
        assert abs(f - ref) <= 1e-6 * (abs(dU / r) + abs(fc)) + 1.6e-7 * abs(qq) / r ** 2


def _minimum_image_sum(s, xyz, chg, rc):
    """O(N^2) pair sum over minimum-image distances < rc (plain numpy)."""
    h, hinv = s.h, np.linalg.inv(s.h)
    n = s.nsites
    ids = s.site_id_array()
    pp = s.potpar
    e = 0.0
    f = np.zeros((3, n))
    for i in range(n - 1):
        d = xyz[:, i + 1:] - xyz[:, i:i + 1]
        sfr = hinv @ d
        sfr -= np.floor(sfr + 0.5)
        d = h @ sfr
        r = np.sqrt((d * d).sum(0))
        m = r < rc
        for jj in np.nonzero(m)[0]:
            j = i + 1 + jj
            p = pp[ids[i], ids[j]]
            u, fij = _pair(s.ptype, r[jj], chg[i] * chg[j], s.control.alpha, p[:units.POT_NPAR[s.ptype]])
            e += u
            f[:, i] -= fij * d[:, jj]
            f[:, j] += fij * d[:, jj]
    return e, f


@pytest.mark.parametrize("mk", ["argon", "water"])
def test_link_cells_equal_minimum_image_sum(mk):
    """Strict cutoff with r_c < L/2: the link-cell sum is the minimum-image
    sum (pairs listed beyond r_c enter at 100 r_c, src/force.c:951-954)."""
    if mk == "argon":
        s = system.argon_fcc(ncell=3, jitter=0.1, cutoff=7.5)
    else:
        s = system.water_box(27, seed=4, cutoff=4.5, alpha=0.6)
    s.control.strict_cutoff = True
    o = oracle(s)
    xyz = o.make_sites()
    r = o.force_calc(xyz)
    e, f = _minimum_image_sum(s, xyz, o.chg, s.control.cutoff)
    e -= r["info"].eintra if mk == "water" else 0.0
    assert abs(r["pe"] - e) <= 1e-9 * abs(e)
    assert np.linalg.norm(r["force"] - f) <= 1e-9 * np.linalg.norm(f)


@pytest.mark.parametrize("case", ["water", "salt", "silica"])
def test_newton_third_law(case):
    s = {"water": S.water, "salt": S.salt, "silica": S.silica}[case]()
    o = oracle(s)
    xyz = o.make_sites()
    fr = o.force_calc(xyz)["force"]
    fk = o.ewald(xyz)["force"]
    assert np.abs(fr.sum(1)).max() <= 1e-10 * np.abs(fr).max() * np.sqrt(s.nsites)
    assert np.abs(fk.sum(1)).max() <= 1e-10 * np.abs(fk).max() * np.sqrt(s.nsites)


def _nacl(n=4, d=1.0):
    a = 2 * d
    basis = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    cells = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    na = ((cells[:, None] + basis[None]) / n).reshape(-1, 3)
    cl = ((cells[:, None] + basis[None] + np.array([.5, 0, 0])) / n).reshape(-1, 3)
    pos = np.concatenate([na, cl]) - 0.5 + 0.25 / n
    pos -= np.floor(pos + 0.5)
    q = np.array([0.0, 1.0, -1.0])
    sp = [system.Species("na", len(na), [1], np.zeros((1, 3))),
          system.Species("cl", len(cl), [2], np.zeros((1, 3)))]
    ctl = system.Control(cutoff=3.9, alpha=1.5, k_cutoff=2 * 1.5 * 6.1)
    return system.MoldySystem(sp, np.array([0, 23.0, 35.5]), q, system.empty_potpar(3), 0,
                              ctl, np.eye(3) * a * n, pos, np.zeros((0, 4)), "nacl")


def test_madelung_constant_rock_salt():
    s = _nacl()
    o = oracle(s)
    xyz = o.make_sites()
    er = o.force_calc(xyz)["pe"]
    ek = o.ewald(xyz)["pe"]
    M = -(er + ek) / (s.nsites / 2)        # E = -(N/2) M q^2 / d, q = 1, d = 1
    assert abs(M - 1.747564594633) < 5e-6


def test_reciprocal_force_is_energy_gradient():
    s = S.water(n=64, seed=5)
    o = oracle(s)
    xyz = o.make_sites()
    f = o.ewald(xyz)["force"]
    rng = np.random.default_rng(1)
    for i in rng.choice(s.nsites, 5, replace=False):
        for c in range(3):
            hstep = 1e-5
            xp, xm = xyz.copy(), xyz.copy()
            xp[c, i] += hstep
            xm[c, i] -= hstep
            dE = (o.ewald(xp)["pe"] - o.ewald(xm)["pe"]) / (2 * hstep)
            assert abs(f[c, i] + dE) <= 2e-6 * np.abs(f).max()


def test_reciprocal_virial_is_strain_derivative():
    """W_ab = -dE/d eps for the strain r -> (1 + eps e_a e_b^T) r (upper
    triangle, so h stays upper triangular as the k enumeration assumes)."""
    s = S.water(n=64, seed=6)
    o = oracle(s)
    xyz = o.make_sites()
    W = o.ewald(xyz)["stress"]
    for a, b in [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]:
        dE = []
        for sgn in (1, -1):
            eps = sgn * 1e-6
            St = np.eye(3)
            St[a, b] += eps
            s2 = s.copy()
            s2.h = St @ s.h
            o2 = oracle(s2)
            dE.append(o2.ewald(St @ xyz)["pe"])
        deriv = (dE[0] - dE[1]) / 2e-6
        assert abs(W[a, b] + deriv) <= 1e-5 * np.abs(W).max()


def test_cellbin_edge_clamp():
    import ctypes as C
    from oracle.oracle import lib
    L = lib()
    eps = 8 * L.orc_precision()
    assert eps == 8 * 2.0 ** -52
    st = C.c_int(0)
    assert L.orc_cellbin(-0.5, 10, 10.0, eps, C.byref(st)) == 0 and st.value == 0
    assert L.orc_cellbin(-0.5 - 0.5 * eps, 10, 10.0, eps, C.byref(st)) == 0 and st.value == 0
    assert L.orc_cellbin(0.5, 10, 10.0, eps, C.byref(st)) == 9 and st.value == 0
    assert L.orc_cellbin(0.5 - 1e-17, 10, 10.0, eps, C.byref(st)) == 9 and st.value == 0
    assert L.orc_cellbin(0.0, 10, 10.0, eps, C.byref(st)) == 5
    L.orc_cellbin(0.5 + 4 * eps, 10, 10.0, eps, C.byref(st))
    assert st.value == 1                       # out of range (force.c:130)


@pytest.mark.parametrize("h", [np.diag([20.0, 22.0, 18.0]),
                               np.array([[19.612, -9.806, 0.0], [0.0, 16.98449, 0.0],
                                         [0.0, 0.0, 21.572]])])
def test_neighbour_offsets_brute_force(h):
    """Cell-centre criterion: half space (ix>0, or ix=0 and iy>0, or ix=iy=0
    and iz>=0) of offsets whose centre distance is below r_c."""
    s = system.water_box(8)
    s.h = h
    s.control.cutoff = 6.5
    o = oracle(s)
    nx, ny, nz = o.choose_cells()
    got = {tuple(x) for x in o.neighbour_offsets(strict=False)}
    want = set()
    for ix in range(0, nx + 1):
        for iy in range(-ny, ny + 1):
            for iz in range(-nz, nz + 1):
                if ix == 0 and (iy < 0 or (iy == 0 and iz < 0)):
                    continue
                v = h @ np.array([ix / nx, iy / ny, iz / nz])
                if v @ v < 6.5 ** 2:
                    want.add((ix, iy, iz))
    # the reference's loops stop short of +m in y and z (force.c:190-192);
    # those cells lie beyond r_c for these boxes, so the sets agree
    assert got == want


@pytest.mark.parametrize("strict", [False, True])
def test_pair_set_is_symmetric_closure_of_cell_lists(strict):
    """Every unordered pair appears once in the site lists; with the strict
    criterion every pair within r_c is listed (the cell-centre criterion may
    miss some, which is why Moldy offers strict-cutoff)."""
    s = S.water(n=64, seed=7)
    s.control.strict_cutoff = strict
    o = oracle(s)
    xyz = o.make_sites()
    pl = o.site_lists(xyz)
    cells = o.fill_cells(xyz)
    nx, ny, nz = cells["dims"]
    seen = {}
    for c in range(nx * ny * nz):
        ent = cells["cell_entries"][cells["cell_start"][c]:cells["cell_start"][c + 1]]
        own = [cells["entry_isite"][e] for e in ent]
        lo, hi = pl["start"][c], pl["start"][c + 1]
        nab = pl["nab"][lo:hi]
        img = pl["pbctrans"][lo:hi]
        for p, i in enumerate(own):
            for j, k in zip(nab[p + 1:], img[p + 1:]):
                key = (min(i, j), max(i, j), k if i < j else 26 - k)
                seen[key] = seen.get(key, 0) + 1
    assert max(seen.values()) == 1
    # count: every site pair within r_c (minimum image) must be present
    h = s.h
    hinv = np.linalg.inv(h)
    if not strict:
        return
    pairs = {(k[0], k[1]) for k in seen}
    for i in range(0, s.nsites, 7):
        d = xyz - xyz[:, i:i + 1]
        f = hinv @ d
        f -= np.floor(f + 0.5)
        r = np.sqrt(((h @ f) ** 2).sum(0))
        for j in np.nonzero((r < s.control.cutoff) & (np.arange(s.nsites) != i))[0]:
            assert (min(i, j), max(i, j)) in pairs, (i, j)
