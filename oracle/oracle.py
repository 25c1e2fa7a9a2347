"""CPU oracle for the D3Q19 MRT channel-flow hot path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check the GPU results.  The product (liblbm.so and its
Python/Fortran hosts) never imports this module.

`OracleChannel` holds the same module state as Channel-Flow/var_inc.f90:107-139
(f, rho, ux, uy, uz, force_realx/y/z) and exposes the reference subroutines
under their own names, each calling the C restatement in lbm_oracle.c:

    collision_MRT   Channel-Flow/collision.f90:24-273
    macrovar        Channel-Flow/collision.f90:378-463
    rhoupdat        Channel-Flow/collision.f90:469-480
    avedensity      Channel-Flow/collision.f90:487-513
    FORCING         Channel-Flow/collision.f90:515-527
    initpop         Channel-Flow/initial.f90:19-46
    initvel         Channel-Flow/initial.f90:75-147
    statistc2       Channel-Flow/saveload.f90:1348-1502
    diag            Channel-Flow/saveload.f90:1507-1676
    vortcalc        Channel-Flow/saveload.f90:3929-4023
    dissipation     Channel-Flow/saveload.f90:4644-4882 (dissipation_calc)

Array layout is the reference's: f(0:18,lx,ly,lz) column-major, held here as a
C-ordered numpy array of shape (lz, ly, lx, 19); macroscopic fields (lz, ly, lx).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# D3Q19_ORACLE_LIB names another build of the same sources; only the test that
# plants mistakes in the oracle (tests/test_oracle_mutations.py) sets it.
_LIB_PATH = os.environ.get("D3Q19_ORACLE_LIB") or os.path.join(_HERE, "liblbm_oracle.so")

_DBL = ctypes.c_double
_INT = ctypes.c_int


class OrcParams(ctypes.Structure):
    """Mirror of orc_params (lbm_oracle.h)."""
    _fields_ = [(n, _INT) for n in ("lx", "ly", "lz", "nx", "mrttype")] + [
        (n, _DBL) for n in (
            "visc", "rstar", "ustar", "force_in_y", "ystar", "force_mag", "tau",
            "s1", "s2", "s4", "s9", "s10", "s13", "s16",
            "omegepsl", "omegepslj", "omegxx",
            "coef1", "coef2", "coef3", "coef4", "coef5", "coef3i", "coef4i",
            "val8",
            "val1i", "val2i", "val3i", "val4i", "val5i", "val6i", "val7i", "val8i", "val9i",
            "ww0", "ww1", "ww2")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def build():
    """Compile liblbm_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(OrcParams)
        D = ctypes.c_void_p
        L.orc_para.argtypes = [P, _INT, _INT, _INT, _INT, _INT]
        L.orc_collide_node.argtypes = [P, D] + [_DBL] * 7 + [D]
        L.orc_collision_mrt.argtypes = [P] + [D] * 8
        L.orc_collision_mrt_ab.argtypes = [P] + [D] * 9 + [_INT]
        L.orc_macrovar.argtypes = [P] + [D] * 8
        L.orc_rhoupdat.argtypes = [P, D, D]
        L.orc_avedensity.argtypes = [P, D]
        L.orc_avedensity.restype = _DBL
        L.orc_forcing.argtypes = [P, D, D, D]
        L.orc_forcingp.argtypes = [P, _INT, D, D, D]
        L.orc_initpop.argtypes = [P] + [D] * 5
        L.orc_initvel.argtypes = [P, D, D, D]
        L.orc_statistc2.argtypes = [P] + [D] * 5
        L.orc_diag.argtypes = [P] + [D] * 5
        L.orc_vortcalc.argtypes = [P] + [D] * 8
        L.orc_dissipation_node.argtypes = [P, D, _DBL, _DBL, _DBL, _DBL, D]
        L.orc_dissipation.argtypes = [P] + [D] * 7
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def para(lx, ly, lz, laminar=False, mrttype=-1):
    """Constants of Channel-Flow/para.f90:59-174 for a (lx, ly, lz) grid."""
    p = OrcParams()
    lib().orc_para(ctypes.byref(p), lx, ly, lz, 1 if laminar else 0, mrttype)
    return p


def collide_node(p, fin, rho, ux, uy, uz, fx, fy, fz):
    """Post-collision values f9(ip)+Fbar(ip)/2 of one node (collision.f90:56-220)."""
    fin = np.ascontiguousarray(fin, dtype=np.float64)
    out = np.empty(19)
    lib().orc_collide_node(ctypes.byref(p), _ptr(fin), rho, ux, uy, uz, fx, fy, fz, _ptr(out))
    return out


class OracleChannel:
    """Module state of the reference channel solver, one MPI rank (periodic y/z)."""

    def __init__(self, lx, ly, lz, laminar=False, mrttype=-1, params=None):
        self.p = params if params is not None else para(lx, ly, lz, laminar, mrttype)
        self.lx, self.ly, self.lz = lx, ly, lz
        shp = (lz, ly, lx)
        self.f = np.zeros(shp + (19,))
        self.rho = np.zeros(shp)
        self.ux = np.zeros(shp)
        self.uy = np.zeros(shp)
        self.uz = np.zeros(shp)
        self.fx = np.zeros(shp)
        self.fy = np.zeros(shp)
        self.fz = np.zeros(shp)

    # --- reference subroutines -------------------------------------------
    def FORCING(self):
        lib().orc_forcing(ctypes.byref(self.p), _ptr(self.fx), _ptr(self.fy), _ptr(self.fz))

    def FORCINGP(self, istep):
        """Perturbation forcing of collision.f90:529-602 at time step istep."""
        lib().orc_forcingp(ctypes.byref(self.p), int(istep), _ptr(self.fx), _ptr(self.fy), _ptr(self.fz))

    def initvel(self):
        lib().orc_initvel(ctypes.byref(self.p), _ptr(self.ux), _ptr(self.uy), _ptr(self.uz))

    def initpop(self):
        lib().orc_initpop(ctypes.byref(self.p), _ptr(self.f), _ptr(self.rho), _ptr(self.ux),
                          _ptr(self.uy), _ptr(self.uz))

    def collision_MRT(self):
        lib().orc_collision_mrt(ctypes.byref(self.p), _ptr(self.f), _ptr(self.rho), _ptr(self.ux),
                                _ptr(self.uy), _ptr(self.uz), _ptr(self.fx), _ptr(self.fy),
                                _ptr(self.fz))

    def collision_MRT_ab(self, nthreads=0):
        """Same step, computed out of place and in parallel (bitwise equal)."""
        out = np.empty_like(self.f)
        lib().orc_collision_mrt_ab(ctypes.byref(self.p), _ptr(self.f), _ptr(out), _ptr(self.rho),
                                   _ptr(self.ux), _ptr(self.uy), _ptr(self.uz), _ptr(self.fx),
                                   _ptr(self.fy), _ptr(self.fz), nthreads)
        self.f = out

    def macrovar(self):
        lib().orc_macrovar(ctypes.byref(self.p), _ptr(self.f), _ptr(self.rho), _ptr(self.ux),
                           _ptr(self.uy), _ptr(self.uz), _ptr(self.fx), _ptr(self.fy), _ptr(self.fz))

    def rhoupdat(self):
        lib().orc_rhoupdat(ctypes.byref(self.p), _ptr(self.f), _ptr(self.rho))

    def avedensity(self):
        return lib().orc_avedensity(ctypes.byref(self.p), _ptr(self.rho))

    def statistc2(self):
        out = np.empty((self.lx, 15))
        lib().orc_statistc2(ctypes.byref(self.p), _ptr(self.rho), _ptr(self.ux), _ptr(self.uy),
                            _ptr(self.uz), _ptr(out))
        return out

    def diag(self):
        out = np.empty(13)
        lib().orc_diag(ctypes.byref(self.p), _ptr(self.rho), _ptr(self.ux), _ptr(self.uy),
                       _ptr(self.uz), _ptr(out))
        return out

    def vortcalc(self, solid=None, omg=None):
        """Channel-Flow/saveload.f90:3929-4023; (ox, oy, oz) of the current fields."""
        out = [np.empty_like(self.ux) for _ in range(3)]
        sp = None if solid is None else np.ascontiguousarray(solid, dtype=np.int32).ctypes.data_as(ctypes.c_void_p)
        op = None if omg is None else _ptr(np.ascontiguousarray(omg, dtype=np.float64))
        lib().orc_vortcalc(ctypes.byref(self.p), _ptr(self.ux), _ptr(self.uy), _ptr(self.uz), sp, op,
                           *(_ptr(a) for a in out))
        return tuple(out)

    def dissipation(self, solid=None):
        """Channel-Flow/saveload.f90:4644-4882; array (lx, 6)."""
        out = np.empty((self.p.lx, 6))
        sp = None if solid is None else np.ascontiguousarray(solid, dtype=np.int32).ctypes.data_as(ctypes.c_void_p)
        lib().orc_dissipation(ctypes.byref(self.p), _ptr(self.f), _ptr(self.rho), _ptr(self.ux), _ptr(self.uy),
                              _ptr(self.uz), sp, _ptr(out))
        return out

    # --- driver sequences (Channel-Flow/main.f90) --------------------------
    def prerelax(self, rhoepsl=1.0e-5, maxiter=15000):
        """Pre-relaxation loop of main.f90:70-90 with u held at its current value.
        Returns (iterations, last max|drho|)."""
        istep = 0
        while True:
            rhop = self.rho.copy()
            self.rhoupdat()
            self.collision_MRT()
            rhoerrmax = float(np.max(np.abs(self.rho - rhop)))
            if rhoerrmax <= rhoepsl or istep > maxiter:
                return istep, rhoerrmax
            istep += 1

    def run_forcingp(self, istep0, nsteps, npforcing):
        """Main loop with the perturbation forcing of main.f90:148-153 switched on:
        FORCINGP before the collision while istep < npforcing, FORCING at npforcing."""
        for istep in range(istep0 + 1, istep0 + nsteps + 1):
            if istep < npforcing:
                self.FORCINGP(istep)
            elif istep == npforcing:
                self.FORCING()
            self.collision_MRT()
            self.macrovar()

    def run(self, nsteps, ab=False, nthreads=0):
        """Main loop body of main.f90:142-167 without I/O: collision_MRT, macrovar."""
        for _ in range(nsteps):
            if ab:
                self.collision_MRT_ab(nthreads)
            else:
                self.collision_MRT()
            self.macrovar()


# --- rigid particles (oracle/particles_oracle.c) ------------------------------

class PGeom(ctypes.Structure):
    """Mirror of orc_pgeom (particles_oracle.h)."""
    _fields_ = [(n, _INT) for n in ("lx", "ly", "lz", "np", "maxlink", "maxfill", "pressure_force",
                                     "reserved")] + [
        (n, _DBL) for n in ("R", "R2", "mass", "inertia", "volume", "rhopart", "mingap", "mingap_w",
                            "stf0", "stf1", "stf0_w", "stf1_w", "visc", "lub_cut", "lub_min", "fx", "fy", "fz")]


class PState(ctypes.Structure):
    """Mirror of orc_pstate (particles_oracle.h)."""
    _fields_ = [(n, ctypes.c_void_p) for n in ("X", "Xold", "V", "W", "TH", "Aprev", "ALprev", "FH", "TQ", "FC",
                                               "FL", "FHprev", "TQprev",
                                               "solid", "link_node", "link_dir", "link_q", "link_count",
                                               "fill_node", "fill_count")] + [("nsteps_done", _INT)]


def _plib():
    L = lib()
    if not hasattr(L, "_p_ready"):
        G, S = ctypes.POINTER(PGeom), ctypes.POINTER(PState)
        L.orc_p_classify.argtypes = [G, S, _INT]
        L.orc_p_classify.restype = _INT
        L.orc_p_ibb.argtypes = [G, S, ctypes.c_void_p]
        L.orc_p_step.argtypes = [ctypes.POINTER(OrcParams), G, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 _DBL, _DBL, _DBL, _DBL]
        L.orc_p_move.argtypes = [G, S]
        L.orc_p_refill.argtypes = [G, S, ctypes.c_void_p]
        L.orc_p_rho_mean.argtypes = [G, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_p_rho_mean.restype = _DBL
        L.orc_p_macrovar.argtypes = [G, S, ctypes.c_void_p, _DBL, _DBL, _DBL, _DBL] + [ctypes.c_void_p] * 4
        L.orc_device_sum.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p]
        L.orc_device_sum.restype = _DBL
        L._p_ready = True
    return L


def particle_geometry(lx, ly, lz, npart, rad, rhopart=1.0, mingap=3.0, mingap_w=3.0, stf0=0.025, stf1=0.002,
                      stf0_w=0.025, stf1_w=0.002, visc=0.0, lub_cut=2.0 / 3.0, lub_min=0.01,
                      pressure_force=1, force=(0.0, 0.0, 0.0), maxlink=0):
    """Particle constants formed exactly as lbm_particles_create forms them."""
    import math
    g = PGeom()
    g.lx, g.ly, g.lz, g.np = lx, ly, lz, npart
    area = 4.0 * math.pi * (rad + 1.0) * (rad + 1.0)
    g.maxlink = maxlink if maxlink > 0 else int(12.0 * area) + 256
    g.maxfill = int(4.0 * area) + 64
    g.pressure_force = pressure_force
    g.R = rad
    g.R2 = rad * rad
    g.volume = 4.0 / 3.0 * math.pi * rad * rad * rad
    g.mass = rhopart * g.volume
    g.inertia = 0.4 * g.mass * rad * rad
    g.rhopart = rhopart
    g.mingap, g.mingap_w = mingap, mingap_w
    g.stf0, g.stf1, g.stf0_w, g.stf1_w = stf0, stf1, stf0_w, stf1_w
    g.visc, g.lub_cut, g.lub_min = visc, lub_cut, lub_min
    g.fx, g.fy, g.fz = force
    return g


class OracleParticleChannel:
    """Particle-laden channel in the pull form the device uses: two post-collision
    lattices L[0], L[1] (19*N each, population-major), node flags, cut links."""

    def __init__(self, params, geom, L0, pos, vel=None, omg=None, L1=None):
        """L0 = None builds a classification-only instance (no lattices)."""
        self.p, self.g = params, geom
        self.N = geom.lx * geom.ly * geom.lz
        np_ = geom.np
        if L0 is None:
            self.L = [np.zeros(1), np.zeros(1)]
        else:
            self.L = [np.ascontiguousarray(L0, dtype=np.float64).reshape(19 * self.N).copy(),
                      np.zeros(19 * self.N) if L1 is None
                      else np.ascontiguousarray(L1, dtype=np.float64).reshape(-1).copy()]
        self.cur = 0
        z3 = lambda: np.zeros(3 * np_)
        self.X = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1).copy()
        self.Xold = self.X.copy()
        self.V = z3() if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1).copy()
        self.W = z3() if omg is None else np.ascontiguousarray(omg, dtype=np.float64).reshape(-1).copy()
        self.TH, self.Aprev, self.ALprev, self.FH, self.TQ, self.FC = (z3() for _ in range(6))
        self.FL, self.FHprev, self.TQprev = (z3() for _ in range(3))
        self.solid = np.full(self.N, -1, dtype=np.int32)
        self.link_node = np.zeros(np_ * geom.maxlink, dtype=np.int32)
        self.link_dir = np.zeros(np_ * geom.maxlink, dtype=np.int32)
        self.link_q = np.zeros(np_ * geom.maxlink)
        self.link_count = np.zeros(np_, dtype=np.int32)
        self.fill_node = np.zeros(np_ * geom.maxfill, dtype=np.int32)
        self.fill_count = np.zeros(np_, dtype=np.int32)
        self.s = PState()
        for name in ("X", "Xold", "V", "W", "TH", "Aprev", "ALprev", "FH", "TQ", "FC", "FL", "FHprev", "TQprev",
                     "solid", "link_node", "link_dir", "link_q", "link_count", "fill_node", "fill_count"):
            setattr(self.s, name, getattr(self, name).ctypes.data_as(ctypes.c_void_p))
        self.s.nsteps_done = 0
        self.force = (geom.fx, geom.fy, geom.fz)
        self.rho_off = 0.0
        self.ave_pending = False
        self.beads_links(first=True)

    def beads_links(self, first=False):
        if _plib().orc_p_classify(ctypes.byref(self.g), ctypes.byref(self.s), 1 if first else 0):
            raise RuntimeError("oracle: particle link or refill list overflow")

    def beads_collision(self):
        _plib().orc_p_ibb(ctypes.byref(self.g), ctypes.byref(self.s), _ptr(self.L[self.cur]))

    def collision_MRT(self):
        _plib().orc_p_step(ctypes.byref(self.p), ctypes.byref(self.g), self.solid.ctypes.data_as(ctypes.c_void_p),
                           _ptr(self.L[self.cur]), _ptr(self.L[1 - self.cur]), *self.force, self.rho_off)
        self.cur = 1 - self.cur
        self.rho_off = 0.0

    def rho_mean(self):
        """avedensity's mean over the fluid nodes of the current lattice (device order)."""
        return _plib().orc_p_rho_mean(ctypes.byref(self.g), self.solid.ctypes.data_as(ctypes.c_void_p),
                                      _ptr(self.L[self.cur]))

    def avedensity(self):
        """Channel-Flow/collision.f90:487-513: the next collision takes rho - mean."""
        self.rho_off = self.rho_mean()
        return self.rho_off

    def macrovar(self):
        """Fields (rho, ux, uy, uz), each (lz, ly, lx), of the current state,
        with the rigid-body velocity and rhopart inside particles."""
        out = [np.empty(self.N) for _ in range(4)]
        _plib().orc_p_macrovar(ctypes.byref(self.g), ctypes.byref(self.s), _ptr(self.L[self.cur]), *self.force,
                               self.rho_off, *(_ptr(a) for a in out))
        return tuple(a.reshape(self.g.lz, self.g.ly, self.g.lx) for a in out)

    def beads_move(self):
        _plib().orc_p_move(ctypes.byref(self.g), ctypes.byref(self.s))

    def beads_filling(self):
        _plib().orc_p_refill(ctypes.byref(self.g), ctypes.byref(self.s), _ptr(self.L[self.cur]))

    def step(self, n=1, istep0=None, nave=100):
        """n steps of the particle-laden main loop: beads_collision, collision_MRT,
        macrovar, beads_move, beads_links, beads_filling.  With istep0 the
        steps are istep0+1 .. istep0+n and, as Channel-Flow/main.f90:163-167
        asks after every step that is a multiple of nave, avedensity's mean
        is taken once the next step's bounce-back has completed the fluid
        nodes' populations, and the next collision subtracts it."""
        for k in range(n):
            self.beads_collision()
            if self.ave_pending:
                self.avedensity()
                self.ave_pending = False
            self.collision_MRT()
            if istep0 is not None and nave > 0 and (istep0 + k + 1) % nave == 0:
                self.ave_pending = True
            self.beads_move()
            self.beads_links()
            self.beads_filling()

    @property
    def post(self):
        return self.L[self.cur]
