"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/_build/liboracle.so (built by ``make -C oracle``) and exposes the
restated reference functions on numpy arrays.  Only tests/, __graft_entry__
.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from moldy_amd import units

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)


class OrcSys(C.Structure):
    _fields_ = [("nsites", C.c_int), ("nmols", C.c_int), ("nspecies", C.c_int),
                ("max_id", C.c_int), ("ptype", C.c_int), ("n_potpar", C.c_int),
                ("spec_nmols", _ip), ("spec_nsites", _ip), ("spec_framework", _ip),
                ("spec_rdof", _ip), ("spec_site_id", _ip), ("spec_p_f_sites", _dp),
                ("site_charge", _dp), ("potpar", _dp),
                ("cutoff", C.c_double), ("subcell", C.c_double), ("alpha", C.c_double),
                ("k_cutoff", C.c_double), ("strict_cutoff", C.c_int), ("molpbc", C.c_int),
                ("surface_dipole", C.c_int)]


class FcInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n_nabors", C.c_int),
                ("n_frame_types", C.c_int), ("n_list_entries", C.c_long),
                ("n_pairs", C.c_long), ("n_too_close", C.c_int),
                ("too_close_i", C.c_int * 16), ("too_close_j", C.c_int * 16),
                ("n_nab_sites", C.c_long), ("eintra", C.c_double)]


class EwInfo(C.Structure):
    _fields_ = [("hmax", C.c_int), ("kmax", C.c_int), ("lmax", C.c_int), ("nhkl", C.c_int),
                ("nsitesxf", C.c_int), ("self_energy", C.c_double),
                ("sheet_energy", C.c_double)]


class OrcDyn(C.Structure):
    _fields_ = [("step", C.c_double), ("ttmass", C.c_double), ("pmass", C.c_double),
                ("gkt", C.c_double), ("const_temp", C.c_int), ("nosymmetric_rot", C.c_int),
                ("spec_mass", C.POINTER(C.c_double)), ("spec_inertia", C.POINTER(C.c_double)),
                ("const_pressure", C.c_int), ("pressure", C.c_double), ("strain_mask", C.c_int)]


class OrcMdState(C.Structure):
    _fields_ = [("c_of_m", C.POINTER(C.c_double)), ("quat", C.POINTER(C.c_double)),
                ("mom", C.POINTER(C.c_double)), ("amom", C.POINTER(C.c_double)),
                ("ts", C.c_double), ("tsmom", C.c_double), ("H_0", C.c_double),
                ("bad_quat", C.c_int), ("h", C.c_double * 9), ("hmom", C.c_double * 9)]


def _build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(build_if_missing=True):
    if not os.path.exists(LIB) and build_if_missing:
        _build()
    lib = C.CDLL(LIB)
    lib.orc_det.restype = C.c_double
    lib.orc_err_fn.restype = C.c_double
    lib.orc_err_fn.argtypes = [C.c_double]
    lib.orc_precision.restype = C.c_double
    lib.orc_poteval.restype = C.c_double
    lib.orc_poteval.argtypes = [_dp, C.c_double, C.c_int, C.c_double, C.c_double]
    lib.orc_dist_pot.restype = C.c_double
    lib.orc_dist_pot.argtypes = [_dp, C.c_double, C.c_int]
    lib.orc_distant_const.restype = C.c_double
    lib.orc_distant_const.argtypes = [C.POINTER(OrcSys), C.c_double, C.c_int]
    lib.orc_site_lists.restype = C.c_long
    lib.orc_cellbin.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, _ip]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Oracle:
    """Oracle bound to one MoldySystem (keeps the flat arrays alive)."""

    def __init__(self, system):
        self.sys = system
        self.f = system.flat()
        f, c = self.f, system.control
        self.cs = OrcSys(system.nsites, system.nmols, system.nspecies, system.max_id,
                         system.ptype, system.n_potpar,
                         _p(f["spec_nmols"], _ip), _p(f["spec_nsites"], _ip),
                         _p(f["spec_framework"], _ip), _p(f["spec_rdof"], _ip),
                         _p(f["spec_site_id"], _ip), _p(f["spec_p_f_sites"]),
                         _p(f["site_charge"]), _p(f["potpar"]),
                         c.cutoff, c.subcell, c.alpha, c.k_cutoff, int(c.strict_cutoff),
                         int(c.molpbc), int(c.surface_dipole))
        self.h = np.ascontiguousarray(system.h, np.float64)
        self.chg = system.site_charges()

    # ------------------------------------------------------------------
    def make_sites(self, molflag=None, c_of_m=None, quat=None):
        s = self.sys
        n = s.nsites
        xyz = np.zeros((3, n))
        mf = (1 if s.control.molpbc else 0) if molflag is None else molflag
        cm = np.ascontiguousarray(s.c_of_m if c_of_m is None else c_of_m, np.float64)
        q = np.ascontiguousarray(self.f["quat"] if quat is None else quat, np.float64)
        lib().orc_make_sites(C.byref(self.cs), _p(self.h), _p(cm), _p(q), mf,
                             _p(xyz[0]), _p(xyz[1]), _p(xyz[2]))
        return xyz

    def choose_cells(self):
        nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
        lib().orc_choose_cells(C.byref(self.cs), _p(self.h), C.byref(nx), C.byref(ny),
                               C.byref(nz))
        return nx.value, ny.value, nz.value

    def fill_cells(self, xyz):
        s = self.sys
        nx, ny, nz = self.choose_cells()
        n = s.nsites + 1
        ec, ei, en, ef, ce = (np.zeros(n, np.int32) for _ in range(5))
        cs = np.zeros(nx * ny * nz + 1, np.int32)
        nft, st = C.c_int(), C.c_int(0)
        ne = lib().orc_fill_cells(C.byref(self.cs), _p(self.h), _p(xyz[0]), _p(xyz[1]),
                                  _p(xyz[2]), _p(np.ascontiguousarray(s.c_of_m)), nx, ny, nz,
                                  _p(ec, _ip), _p(ei, _ip), _p(en, _ip), _p(ef, _ip),
                                  _p(cs, _ip), _p(ce, _ip), C.byref(nft), C.byref(st))
        return dict(ne=ne, entry_cell=ec[:ne], entry_isite=ei[:ne], entry_num=en[:ne],
                    entry_ftype=ef[:ne], cell_start=cs, cell_entries=ce[:ne],
                    n_frame_types=nft.value, status=st.value, dims=(nx, ny, nz))

    def neighbour_offsets(self, strict=None, cutoff=None, dims=None):
        c = self.sys.control
        nx, ny, nz = dims or self.choose_cells()
        cap = 1 << 20
        off = np.zeros((cap, 3), np.int32)
        n = lib().orc_neighbour_list(_p(self.h), C.c_double(c.cutoff if cutoff is None else cutoff),
                                     nx, ny, nz, int(c.strict_cutoff if strict is None else strict),
                                     _p(off, _ip), cap)
        if n < 0:
            raise RuntimeError(f"neighbour_list error {n}")
        return off[:n].copy()

    def site_lists(self, xyz, strict=None, cutoff=None):
        c = self.sys.control
        nx, ny, nz = self.choose_cells()
        ncell = nx * ny * nz
        starts = np.zeros(ncell + 1, np.int64)
        st = C.c_int(0)
        strict = int(c.strict_cutoff if strict is None else strict)
        cut = C.c_double(c.cutoff if cutoff is None else cutoff)
        cm = np.ascontiguousarray(self.sys.c_of_m, np.float64)
        total = lib().orc_site_lists(C.byref(self.cs), _p(self.h), _p(xyz[0]), _p(xyz[1]),
                                     _p(xyz[2]), _p(cm), strict, cut, _p(starts, _lp),
                                     None, None, None, C.c_long(0), C.byref(st))
        if total < 0:
            raise RuntimeError(f"site_lists error {total}")
        nab = np.zeros(total + 1, np.int32)
        pbc = np.zeros(total + 1, np.int32)
        nf = np.zeros((ncell, 2), np.int32)
        lib().orc_site_lists(C.byref(self.cs), _p(self.h), _p(xyz[0]), _p(xyz[1]), _p(xyz[2]),
                             _p(cm), strict, cut, _p(starts, _lp), _p(nab, _ip), _p(pbc, _ip),
                             _p(nf, _ip), C.c_long(total + 1), C.byref(st))
        return dict(start=starts, nab=nab[:total], pbctrans=pbc[:total], nfnab=nf,
                    status=st.value)

    def force_calc(self, xyz, fxyz=None):
        n = self.sys.nsites
        f = np.zeros((3, n)) if fxyz is None else fxyz
        pe = C.c_double(0.0)
        stress = np.zeros(9)
        info = FcInfo()
        cm = np.ascontiguousarray(self.sys.c_of_m, np.float64)
        xyz = np.ascontiguousarray(xyz)
        rc = lib().orc_force_calc(C.byref(self.cs), _p(self.h), _p(xyz[0]), _p(xyz[1]),
                                  _p(xyz[2]), _p(cm), _p(self.chg), _p(f[0]), _p(f[1]),
                                  _p(f[2]), C.byref(pe), _p(stress), C.byref(info))
        return dict(rc=rc, force=f, pe=pe.value, stress=stress.reshape(3, 3), info=info)

    def ewald(self, xyz, fxyz=None):
        n = self.sys.nsites
        f = np.zeros((3, n)) if fxyz is None else fxyz
        pe = C.c_double(0.0)
        stress = np.zeros(9)
        info = EwInfo()
        xyz = np.ascontiguousarray(xyz)
        rc = lib().orc_ewald(C.byref(self.cs), _p(self.h), _p(xyz[0]), _p(xyz[1]), _p(xyz[2]),
                             _p(self.chg), _p(f[0]), _p(f[1]), _p(f[2]), C.byref(pe),
                             _p(stress), C.byref(info))
        return dict(rc=rc, force=f, pe=pe.value, stress=stress.reshape(3, 3), info=info)

    def kset(self, k_cutoff=None):
        kc = self.sys.control.k_cutoff if k_cutoff is None else k_cutoff
        hm, km, lm = C.c_int(), C.c_int(), C.c_int()
        n = lib().orc_kset(_p(self.h), C.c_double(kc), None, None, 0, C.byref(hm),
                           C.byref(km), C.byref(lm))
        hkl = np.zeros((n + 1, 3), np.int32)
        kv = np.zeros((n + 1, 3))
        lib().orc_kset(_p(self.h), C.c_double(kc), _p(hkl, _ip), _p(kv), n + 1, None, None, None)
        return hkl[:n], kv[:n], (hm.value, km.value, lm.value)

    def eval_forces(self, c_of_m=None, quat=None):
        s = self.sys
        cm = np.ascontiguousarray(s.c_of_m if c_of_m is None else c_of_m, np.float64)
        q = np.ascontiguousarray(self.f["quat"] if quat is None else quat, np.float64)
        pe = np.zeros(2)
        dip = np.zeros(3)
        stress = np.zeros(9)
        force = np.zeros((s.nmols, 3))
        torque = np.zeros((max(s.nmols_r, 1), 3))
        sxyz = np.zeros((3, s.nsites))
        sf = np.zeros((3, s.nsites))
        fci, ewi = FcInfo(), EwInfo()
        rc = lib().orc_eval_forces(C.byref(self.cs), _p(self.h), _p(cm), _p(q), _p(pe),
                                   _p(dip), _p(stress), _p(force), _p(torque), _p(sxyz),
                                   _p(sf), C.byref(fci), C.byref(ewi))
        return dict(rc=rc, pe=pe, dip=dip, stress=stress.reshape(3, 3), force=force,
                    torque=torque[:s.nmols_r], site_xyz=sxyz, site_force=sf, fc=fci, ew=ewi)

    def md(self, dyn, mom, amom, c_of_m=None, quat=None, ts=1.0, tsmom=0.0, H_0=0.0):
        """Oracle MD run: returns a callable step(init_H_0) -> (rc, pe, stress,
        rec) that advances its own copy of the state, and the state dict.
        rec: KE_trans, KE_rot, pe0, pe1, s, p_s, H_0, det(h), hmom[9], h[9] at the
        H_0 point of do_step.  The cell starts at the system's h with zero
        momentum."""
        s = self.sys
        mass, inertia = s.species_mass_inertia()
        st = dict(c_of_m=np.array(s.c_of_m if c_of_m is None else c_of_m, np.float64),
                  quat=np.array(self.f["quat"] if quat is None else quat, np.float64).reshape(-1, 4),
                  mom=np.array(mom, np.float64).reshape(-1, 3),
                  amom=np.array(amom, np.float64).reshape(-1, 4),
                  mass=np.ascontiguousarray(mass), inertia=np.ascontiguousarray(inertia))
        d = OrcDyn(dyn.step, dyn.ttmass, dyn.pmass,
                   s.degrees_of_freedom() * units.KB * dyn.temp, int(dyn.const_temp),
                   int(dyn.nosymmetric_rot), _p(st["mass"]), _p(st["inertia"]),
                   int(dyn.const_pressure), float(dyn.pressure), int(dyn.strain_mask))
        ms = OrcMdState(_p(st["c_of_m"]), _p(st["quat"]), _p(st["mom"]), _p(st["amom"]),
                        ts, tsmom, H_0, -1)
        for i in range(9):
            ms.h[i] = float(np.asarray(self.h).reshape(-1)[i])
            ms.hmom[i] = 0.0
        st["_keep"] = (d, ms)

        def step(init_H_0=False):
            pe, stress, rec = np.zeros(2), np.zeros(9), np.zeros(26)
            rc = lib().orc_do_step(C.byref(self.cs), C.byref(d), C.byref(ms),
                                   int(init_H_0), _p(pe), _p(stress), _p(rec))
            st["ts"], st["tsmom"], st["H_0"], st["bad_quat"] = ms.ts, ms.tsmom, ms.H_0, ms.bad_quat
            st["h"] = np.array(ms.h[:]).reshape(3, 3)
            st["hmom"] = np.array(ms.hmom[:]).reshape(3, 3)
            return rc, pe, stress.reshape(3, 3), rec
        return step, st

    def rdf(self, xyz, limit, nbins, rdf=None, c_of_m=None):
        """rdf_inner/rdf_accum: float bins [(max_id-1)*max_id/2][nbins]."""
        s = self.sys
        nblk = (s.max_id - 1) * s.max_id // 2
        out = np.zeros((nblk, nbins), np.float32) if rdf is None else rdf
        xyz = np.ascontiguousarray(xyz, np.float64)
        cm = np.ascontiguousarray(s.c_of_m if c_of_m is None else c_of_m, np.float64)
        rc = lib().orc_rdf(C.byref(self.cs), _p(self.h), _p(xyz[0]), _p(xyz[1]), _p(xyz[2]),
                           _p(cm), C.c_double(limit), int(nbins),
                           out.ctypes.data_as(C.POINTER(C.c_float)))
        return rc, out

    def distant_const(self, iflag):
        return lib().orc_distant_const(C.byref(self.cs), C.c_double(self.sys.control.cutoff),
                                       iflag)

    def poteval(self, params, r, chgsq=0.0):
        p = np.zeros(8)
        p[:len(params)] = params
        return lib().orc_poteval(_p(p), C.c_double(r), self.sys.ptype, C.c_double(chgsq),
                                 C.c_double(self.sys.control.alpha))
