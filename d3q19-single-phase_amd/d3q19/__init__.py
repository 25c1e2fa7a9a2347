"""MI355X D3Q19 MRT channel-flow core: Python host over liblbm's C ABI (include/lbm.h).

`Channel` mirrors the reference's channel solver one rank at a time: its
methods carry the names of the Fortran subroutines they replace
(Channel-Flow/collision.f90: collision_MRT, macrovar, rhoupdat, avedensity,
FORCING; Channel-Flow/initial.f90: initvel, initpop) and act on a lattice
that lives on the GPU.  Host arrays use the reference's layout: f(0:18,lx,ly,lz)
held as a C-ordered numpy array of shape (lz, ly, lx, 19), fields as
(lz, ly, lx).

There is no CPU fallback: importing this package requires the built
liblbm.so, and every call runs on the GPU.
"""
import ctypes
import os

import numpy as np

from .params import para, loglaw_profile, CIX, CIY, CIZ, IPOPP  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm.so")

FP64, FP32, FP16, FP16C = 0, 1, 2, 3
_PREC = {"fp64": FP64, "f64": FP64, "fp32": FP32, "f32": FP32, "fp16": FP16, "f16": FP16, "fp16c": FP16C}
STORAGE_BYTES = {FP64: 8, FP32: 4, FP16: 2, FP16C: 2}


class LbmError(RuntimeError):
    """A liblbm call returned an error status."""

    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed (status {code}): {msg}")
        self.code = code


class LbmConfig(ctypes.Structure):
    """Mirror of struct lbm_config in include/lbm.h."""
    _fields_ = [(n, ctypes.c_int) for n in ("lx", "ly", "lz", "precision", "device", "reserved")] + [
        (n, ctypes.c_double) for n in (
            "s1", "s2", "s4", "s9", "s10", "s13", "s16",
            "omegepsl", "omegepslj", "omegxx",
            "coef1", "coef2", "coef3", "coef4", "coef5", "coef3i", "coef4i",
            "val8",
            "val1i", "val2i", "val3i", "val4i", "val5i", "val6i", "val7i", "val8i", "val9i",
            "ww0", "ww1", "ww2")]


class ParticleConfig(ctypes.Structure):
    """Mirror of struct lbm_particle_config in include/lbm.h."""
    _fields_ = [("npart", ctypes.c_int), ("maxlink", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in ("rad", "rhopart", "mingap", "mingap_w", "stf0", "stf1", "stf0_w", "stf1_w",
                                       "visc", "lub_cut", "lub_min")] + [
        ("pressure_force", ctypes.c_int), ("nave", ctypes.c_int)]


# Every function the C ABI exports, with its argument types.
_D = ctypes.c_void_p
_H = ctypes.c_void_p
_SIGS = {
    "lbm_create": [ctypes.POINTER(LbmConfig), ctypes.POINTER(ctypes.c_void_p)],
    "lbm_destroy": [_H],
    "lbm_upload_f": [_H, _D],
    "lbm_download_f": [_H, _D],
    "lbm_set_force": [_H, ctypes.c_double, ctypes.c_double, ctypes.c_double],
    "lbm_forcingp": [_H, ctypes.c_int, ctypes.c_double],
    "lbm_run_forcingp": [_H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double],
    "lbm_set_macros": [_H, _D, _D, _D, _D],
    "lbm_initpop": [_H],
    "lbm_set_velocity_separable": [_H, ctypes.c_int, _D, _D, _D],
    "lbm_collision_mrt": [_H],
    "lbm_macrovar": [_H, _D, _D, _D, _D],
    "lbm_rhoupdat": [_H, ctypes.POINTER(ctypes.c_double)],
    "lbm_avedensity": [_H, ctypes.POINTER(ctypes.c_double)],
    "lbm_run": [_H, ctypes.c_int],
    "lbm_mass_sum": [_H, ctypes.POINTER(ctypes.c_double)],
    "lbm_mass_sum_noalloc": [_H, _D],
    "lbm_statistc2": [_H, ctypes.c_double, ctypes.c_double, _D],
    "lbm_diag": [_H, ctypes.c_double, _D],
    "lbm_vortcalc": [_H, _D, _D, _D],
    "lbm_dissipation": [_H, ctypes.c_double, ctypes.c_double, _D],
    "lbm_stats_accumulate": [_H],
    "lbm_stats_read": [_H, ctypes.c_double, ctypes.c_double, _D, ctypes.POINTER(ctypes.c_longlong)],
    "lbm_stats_reset": [_H],
    "lbm_stats_accumulate_filtered": [_H],
    "lbm_stats_read_filtered": [_H, ctypes.c_double, ctypes.c_double, _D, ctypes.POINTER(ctypes.c_longlong)],
    "lbm_sync": [_H],
    "lbm_time_steps": [_H, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "lbm_device_bytes": [_H, ctypes.POINTER(ctypes.c_size_t)],
    "lbm_block_info": [_H, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "lbm_block_rows": [_H, ctypes.POINTER(ctypes.c_int)],
    "lbm_download_post": [_H, _D],
    "lbm_download_post_range": [_H, ctypes.c_longlong, ctypes.c_longlong, _D],
    "lbm_particles_create": [_H, _D, _D, _D, _D],
    "lbm_beads_links": [_H],
    "lbm_beads_filling": [_H],
    "lbm_beads_collision": [_H],
    "lbm_beads_move": [_H],
    "lbm_run_particles": [_H, ctypes.c_int, ctypes.c_int],
    "lbm_particles_get": [_H, _D, _D, _D, _D, _D, _D, _D],
    "lbm_particles_links": [_H, ctypes.POINTER(ctypes.c_int), _D, _D, _D, _D],
    "lbm_particles_solid": [_H, _D],
    "lbm_closure_set": [_H, _D, ctypes.c_double],
    "lbm_closure_clear": [_H],
    "lbm_closure_inputs": [_H, _D],
    "lbm_closure_nut": [_H, _D],
    "lbm_trainer_create": [ctypes.c_int, ctypes.c_longlong, _D, ctypes.POINTER(ctypes.c_void_p)],
    "lbm_trainer_destroy": [_H],
    "lbm_trainer_set_target": [_H, ctypes.c_int],
    "lbm_trainer_add_filtered": [_H, _H, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.POINTER(ctypes.c_longlong)],
    "lbm_trainer_add_samples": [_H, _D, _D, ctypes.c_longlong],
    "lbm_trainer_set_plane_targets": [_H, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, _D],
    "lbm_trainer_samples": [_H, ctypes.POINTER(ctypes.c_longlong), _D, _D],
    "lbm_trainer_grad": [_H, ctypes.c_longlong, ctypes.c_longlong, _D, ctypes.POINTER(ctypes.c_float)],
    "lbm_trainer_train": [_H, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, ctypes.c_longlong, ctypes.c_float,
                          ctypes.POINTER(ctypes.c_float)],
    "lbm_trainer_predict": [_H, ctypes.c_longlong, ctypes.c_longlong, _D],
    "lbm_trainer_params": [_H, _D],
}

MLP_NPARAM, MLP_TOTAL, MLP_IN, MLP_H = 1441, 1453, 10, 32

_lib = None


def load_library(path=LIB_PATH):
    """Load liblbm.so; raise if it has not been built (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"liblbm.so not found at {path}: build it first "
                          "(python -c 'import __graft_entry__ as g; g.build()' or make -C d3q19-single-phase_amd)")
    L = ctypes.CDLL(path)
    for name, argtypes in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    L.lbm_last_error.argtypes = []
    L.lbm_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def exported_symbols():
    return sorted(list(_SIGS) + ["lbm_last_error"])


def _call(name, *args):
    L = load_library()
    rc = getattr(L, name)(*args)
    if rc != 0:
        raise LbmError(rc, name, L.lbm_last_error().decode(errors="replace"))


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "arrays must be C-contiguous float64"
    return a.ctypes.data_as(ctypes.c_void_p)


def _fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous, "arrays must be C-contiguous float32"
    return a.ctypes.data_as(ctypes.c_void_p)


def closure_params(seed=0, scales=None, yscale=100.0, nu=0.0018, init="small", nut_min=0.0):
    """Parameter vector of the learned closure (include/lbm.h: W1, b1, W2, b2, W3, b3,
    9 input scales, y+ divisor, nu, floor m of nu_t/nu = softplus(o) + m).
    init="small": Xavier-uniform weights from SplitMix64 stream `seed`, output
    bias -4 (nu_t ~ 0.018 nu at start) for m = 0, or softplus(b3) = -m
    (nu_t ~ 0 at start) for a negative floor."""
    if not -1.0 < nut_min <= 0.0:
        raise ValueError(f"nut_min {nut_min} outside (-1, 0]")
    from .workloads import uniform
    p = np.zeros(MLP_TOTAL, dtype=np.float32)
    r = uniform(1000 + seed, MLP_NPARAM)
    o = 0

    def take(n, fan_in, fan_out):
        nonlocal o
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        v = (2.0 * r[o:o + n] - 1.0) * lim
        o += n
        return v

    if init == "small":
        p[0:320] = take(320, 10, 32)
        p[352:352 + 1024] = take(1024, 32, 32)
        p[1408:1440] = take(32, 32, 1)
        p[1440] = -4.0 if nut_min == 0.0 else np.log(np.expm1(-nut_min))
    p[1441:1450] = 1.0 if scales is None else np.asarray(scales, dtype=np.float32)
    p[1450] = yscale
    p[1451] = nu
    p[1452] = nut_min
    return p


class Channel:
    """One rank's channel block on one GPU (reference axes: x wall-normal, y streamwise, z spanwise)."""

    def __init__(self, lx, ly, lz, precision="fp64", laminar=False, mrttype=None, device=0, params=None,
                 visc=None, rstar=None):
        self.lx, self.ly, self.lz = int(lx), int(ly), int(lz)
        self.N = self.lx * self.ly * self.lz
        self.prec = _PREC[precision] if isinstance(precision, str) else int(precision)
        self.p = dict(params) if params is not None else para(lx, laminar, mrttype, visc=visc, rstar=rstar)
        cfg = LbmConfig()
        cfg.lx, cfg.ly, cfg.lz = self.lx, self.ly, self.lz
        cfg.precision, cfg.device, cfg.reserved = self.prec, int(device), 0
        for name, _ in LbmConfig._fields_[6:]:
            setattr(cfg, name, float(self.p[name]))
        self._cfg = cfg
        h = ctypes.c_void_p()
        _call("lbm_create", ctypes.byref(cfg), ctypes.byref(h))
        self._h = h

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _call("lbm_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference subroutines --------------------------------------------------
    def FORCING(self):
        """Channel-Flow/collision.f90:515-527: uniform force (0, force_in_y*force_mag, 0)."""
        self.set_force(0.0, self.p["force_in_y"] * self.p["force_mag"], 0.0)

    def set_force(self, fx, fy, fz):
        _call("lbm_set_force", self._h, float(fx), float(fy), float(fz))

    def FORCINGP(self, istep, force_in_y=None):
        """Channel-Flow/collision.f90:529-602: mean force plus the banded
        divergence-free perturbation of time step istep (force_in_y of
        para.f90:65 unless given)."""
        fiy = self.p["force_in_y"] if force_in_y is None else force_in_y
        _call("lbm_forcingp", self._h, int(istep), float(fiy))

    def run_forcingp(self, istep0, nsteps, npforcing, force_in_y=None):
        """Main loop of Channel-Flow/main.f90:142-161 with the perturbation
        forcing switched on (the commented block at :148-153): FORCINGP before
        the collision while istep < npforcing, FORCING at istep == npforcing.
        One library call (lbm_run_forcingp), pairs of steps as blocked sweeps."""
        fiy = self.p["force_in_y"] if force_in_y is None else force_in_y
        _call("lbm_run_forcingp", self._h, int(istep0), int(nsteps), int(npforcing), float(fiy),
              float(self.p["force_in_y"] * self.p["force_mag"]))

    def run_forcingp_stepwise(self, istep0, nsteps, npforcing, force_in_y=None):
        """The same loop as separate FORCINGP / FORCING, collision_MRT and
        macrovar calls, one step at a time (the reference's call sequence)."""
        for istep in range(istep0 + 1, istep0 + nsteps + 1):
            if istep < npforcing:
                self.FORCINGP(istep, force_in_y)
            elif istep == npforcing:
                self.FORCING()
            self.collision_MRT()
            self.macrovar()

    def initvel(self):
        """Channel-Flow/initial.f90:94-115: log-law streamwise profile, u_x = u_z = 0."""
        prof = np.asarray(loglaw_profile(self.lx, self.p["ustar"], self.p["ystar"]))
        uy = np.ascontiguousarray(np.broadcast_to(prof, (self.lz, self.ly, self.lx)))
        zero = np.zeros((self.lz, self.ly, self.lx))
        self.set_macros(ux=zero, uy=uy, uz=zero)

    def set_macros(self, rho=None, ux=None, uy=None, uz=None):
        """Assign the macroscopic field arrays the next collision reads."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (rho, ux, uy, uz)]
        for a in arrs:
            if a is not None and a.size != self.N:
                raise ValueError(f"field of {a.size} values for {self.N} nodes")
        _call("lbm_set_macros", self._h, *[_ptr(a) for a in arrs])

    def set_velocity_separable(self, terms):
        """rho = 0 and u_c = sum of plane (lz, ly) x profile (lx) terms, summed on
        the device in the order given; terms = [(component, plane, profile), ...]."""
        comp = np.ascontiguousarray([int(c) for c, _, _ in terms], dtype=np.int32)
        planes = np.ascontiguousarray(np.stack([np.asarray(p, dtype=np.float64).reshape(self.lz, self.ly)
                                                for _, p, _ in terms]))
        profs = np.ascontiguousarray(np.stack([np.asarray(q, dtype=np.float64).reshape(self.lx)
                                               for _, _, q in terms]))
        _call("lbm_set_velocity_separable", self._h, len(terms), comp.ctypes.data_as(ctypes.c_void_p),
              _ptr(planes), _ptr(profs))

    def initpop(self):
        """Channel-Flow/initial.f90:19-46: rho = 0, f = equilibrium of the current u."""
        _call("lbm_initpop", self._h)

    def collision_MRT(self):
        """Channel-Flow/collision.f90:24-273 (with collisionExchnge as the periodic wrap)."""
        _call("lbm_collision_mrt", self._h)

    def macrovar(self, download=False, out=None):
        """Channel-Flow/collision.f90:378-463.  With download=True returns (rho, ux, uy, uz),
        filling the four (lz, ly, lx) float64 arrays of `out` when given."""
        if not download:
            _call("lbm_macrovar", self._h, None, None, None, None)
            return None
        if out is None:
            out = [np.empty((self.lz, self.ly, self.lx)) for _ in range(4)]
        for a in out:
            if a.shape != (self.lz, self.ly, self.lx) or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("macrovar: out needs four C-contiguous (lz, ly, lx) float64 arrays")
        _call("lbm_macrovar", self._h, *[_ptr(a) for a in out])
        return tuple(out)

    def rhoupdat(self):
        """Channel-Flow/collision.f90:469-480; returns max|rho_new - rho_old| (main.f90:79)."""
        v = ctypes.c_double()
        _call("lbm_rhoupdat", self._h, ctypes.byref(v))
        return v.value

    def avedensity(self):
        """Channel-Flow/collision.f90:487-513; returns the subtracted mean."""
        v = ctypes.c_double()
        _call("lbm_avedensity", self._h, ctypes.byref(v))
        return v.value

    # -- data movement ---------------------------------------------------------
    def upload_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.size != 19 * self.N:
            raise ValueError(f"f has {f.size} values, expected 19*{self.N}")
        _call("lbm_upload_f", self._h, _ptr(f))

    def download_f(self):
        f = np.empty((self.lz, self.ly, self.lx, 19))
        _call("lbm_download_f", self._h, _ptr(f))
        return f

    # -- driver helpers ----------------------------------------------------------
    def run(self, nsteps):
        """nsteps of the main loop body (main.f90:157-161: collision_MRT, macrovar)."""
        _call("lbm_run", self._h, int(nsteps))

    def prerelax(self, rhoepsl=1.0e-5, maxiter=15000):
        """Pre-relaxation loop of Channel-Flow/main.f90:70-90 (u held fixed).
        Returns (iterations, last max|drho|)."""
        istep = 0
        while True:
            err = self.rhoupdat()
            self.collision_MRT()
            if err <= rhoepsl or istep > maxiter:
                return istep, err
            istep += 1

    def mass(self):
        v = ctypes.c_double()
        _call("lbm_mass_sum", self._h, ctypes.byref(v))
        return v.value

    def mass_noalloc(self, mass_dev_ptr):
        """lbm_mass_sum_noalloc: enqueue the mass sum into one double of device
        memory at address mass_dev_ptr (no allocation, no synchronisation)."""
        _call("lbm_mass_sum_noalloc", self._h, ctypes.c_void_p(int(mass_dev_ptr)))

    def statistc2(self):
        """Plane profiles of Channel-Flow/saveload.f90:1348-1502; array (lx, 15)."""
        out = np.empty((self.lx, 15))
        _call("lbm_statistc2", self._h, self.p["ustar"], self.p["ystar"], _ptr(out))
        return out

    def diag(self):
        """Run monitor of Channel-Flow/saveload.f90:1507-1676 on the device:
        vmax, imout, jmout, kmout, umean, vmean, wmean, urms, vrms, wrms, volf,
        rhomax, rhomin (means and rms / u*)."""
        out = np.empty(13)
        _call("lbm_diag", self._h, self.p["ustar"], _ptr(out))
        return out

    def vortcalc(self):
        """Vorticity (ox, oy, oz), each (lz, ly, lx), of Channel-Flow/saveload.f90:3929-4023."""
        out = [np.empty((self.lz, self.ly, self.lx)) for _ in range(3)]
        _call("lbm_vortcalc", self._h, *[_ptr(a) for a in out])
        return tuple(out)

    def dissipation(self):
        """dissipation_calc of Channel-Flow/saveload.f90:4644-4882; array (lx, 6)."""
        out = np.empty((self.lx, 6))
        _call("lbm_dissipation", self._h, self.p["tau"], self.p["ystar"], _ptr(out))
        return out

    def stats_accumulate(self):
        _call("lbm_stats_accumulate", self._h)

    def stats_read(self):
        out = np.empty((self.lx, 15))
        n = ctypes.c_longlong()
        _call("lbm_stats_read", self._h, self.p["ustar"], self.p["ystar"], _ptr(out), ctypes.byref(n))
        return out, n.value

    def stats_reset(self):
        """Clear the running (and the filtered running) statistics."""
        _call("lbm_stats_reset", self._h)

    def stats_accumulate_filtered(self):
        """Add the plane sums of the 2x2x2 box-filtered fields."""
        _call("lbm_stats_accumulate_filtered", self._h)

    def stats_read_filtered(self):
        """statistc2 columns of the filtered running average on the grid of
        half the size (ystar of that grid = ystar / 2; u* unchanged under the
        acoustic scaling dx, dt -> 2 dx, 2 dt); array (lx/2, 15) and the
        sample count."""
        out = np.empty((self.lx // 2, 15))
        n = ctypes.c_longlong()
        _call("lbm_stats_read_filtered", self._h, self.p["ustar"], self.p["ystar"] / 2.0, _ptr(out),
              ctypes.byref(n))
        return out, n.value

    def sync(self):
        _call("lbm_sync", self._h)

    def time_steps(self, nsteps):
        """Run nsteps, returning (mean kernel ms, total event ms) from HIP events."""
        a, t = ctypes.c_double(), ctypes.c_double()
        _call("lbm_time_steps", self._h, int(nsteps), ctypes.byref(a), ctypes.byref(t))
        return a.value, t.value

    def device_bytes(self):
        v = ctypes.c_size_t()
        _call("lbm_device_bytes", self._h, ctypes.byref(v))
        return v.value

    def block_info(self):
        """The blocked sweep run() uses for pairs of steps: (axis 'z'|'y', slabs per chunk; 0 = off)."""
        a, k = ctypes.c_int(), ctypes.c_int()
        _call("lbm_block_info", self._h, ctypes.byref(a), ctypes.byref(k))
        return chr(a.value), k.value

    def block_rows(self):
        """Rows per window of the blocked sweep's slabs (0: whole slabs)."""
        w = ctypes.c_int()
        _call("lbm_block_rows", self._h, ctypes.byref(w))
        return w.value

    def download_post(self):
        """Stored post-collision lattice, FP64, shape (19, N) (node = ix + lx*(iy + ly*iz))."""
        out = np.empty((19, self.N))
        _call("lbm_download_post", self._h, _ptr(out))
        return out

    def download_post_range(self, first, count):
        """Nodes [first, first + count) of the stored post-collision lattice, FP64, shape (19, count)."""
        out = np.empty((19, int(count)))
        _call("lbm_download_post_range", self._h, int(first), int(count), _ptr(out))
        return out

    # -- rigid particles (beads_* of Channel-Flow/var_inc.f90:166-168) ----------
    def particles_create(self, pos, vel=None, omg=None, rad=5.0, rhopart=1.0, mingap=3.0, mingap_w=3.0,
                         stf0=0.025, stf1=0.002, stf0_w=0.025, stf1_w=0.002, visc=0.0, lub_cut=2.0 / 3.0,
                         lub_min=0.01, pressure_force=True, nave=100, maxlink=0):
        """Spheres at pos (npart, 3) in node coordinates (node ix at ix - 0.5, 1-based).
        Contact: Feng & Michaelides (2005) eq. 28 with the stiffness pairs
        stf0/stf1 (particles) and stf0_w/stf1_w (walls) of Channel-Flow/
        para.f90:357-360.  Lubrication (Nguyen & Ladd 2002) when visc > 0.
        run_particles takes avedensity after every nave-th step."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        pc = ParticleConfig()
        pc.npart, pc.maxlink = pos.shape[0], int(maxlink)
        pc.rad, pc.rhopart, pc.mingap, pc.mingap_w = rad, rhopart, mingap, mingap_w
        pc.stf0, pc.stf1, pc.stf0_w, pc.stf1_w = stf0, stf1, stf0_w, stf1_w
        pc.visc, pc.lub_cut, pc.lub_min = visc, lub_cut, lub_min
        pc.pressure_force, pc.nave = int(bool(pressure_force)), int(nave)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        w = None if omg is None else np.ascontiguousarray(omg, dtype=np.float64).reshape(-1, 3)
        _call("lbm_particles_create", self._h, ctypes.byref(pc), _ptr(pos), _ptr(v), _ptr(w))
        self.npart = pos.shape[0]
        self.istep = 0

    def beads_links(self):
        _call("lbm_beads_links", self._h)

    def beads_filling(self):
        _call("lbm_beads_filling", self._h)

    def beads_collision(self):
        _call("lbm_beads_collision", self._h)

    def beads_move(self):
        _call("lbm_beads_move", self._h)

    def run_particles(self, nsteps, istep0=None):
        """Steps istep0+1 .. istep0+nsteps of the particle-laden loop (istep0
        defaults to the steps this object has run since particles_create)."""
        if istep0 is None:
            istep0 = self.istep
        _call("lbm_run_particles", self._h, int(istep0), int(nsteps))
        self.istep = int(istep0) + int(nsteps)

    def particles_get(self):
        keys = ("pos", "vel", "omg", "force", "torque", "contact", "lub")
        out = {k: np.empty((self.npart, 3)) for k in keys}
        _call("lbm_particles_get", self._h, *[_ptr(out[k]) for k in keys])
        return out

    def particles_links(self):
        m = ctypes.c_int()
        _call("lbm_particles_links", self._h, ctypes.byref(m), None, None, None, None)
        cap = m.value
        counts = np.empty(self.npart, dtype=np.int32)
        node = np.empty(self.npart * cap, dtype=np.int32)
        dirs = np.empty(self.npart * cap, dtype=np.int32)
        q = np.empty(self.npart * cap)
        _call("lbm_particles_links", self._h, ctypes.byref(m), counts.ctypes.data_as(ctypes.c_void_p),
              node.ctypes.data_as(ctypes.c_void_p), dirs.ctypes.data_as(ctypes.c_void_p), _ptr(q))
        return cap, counts, node, dirs, q

    # -- learned closure ------------------------------------------------------
    def closure_set(self, params, ystar=None):
        """Evaluate the learned closure inside every collision (FP32/FP16 storage)."""
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == MLP_TOTAL
        _call("lbm_closure_set", self._h, _fptr(params), float(self.p["ystar"] if ystar is None else ystar))

    def closure_clear(self):
        _call("lbm_closure_clear", self._h)

    def closure_inputs(self):
        """The ten scaled network inputs of every node, (lz, ly, lx, 10) float32."""
        out = np.empty((self.lz, self.ly, self.lx, 10), dtype=np.float32)
        _call("lbm_closure_inputs", self._h, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def closure_nut(self):
        """nu_t / nu the fused collision applies at every node, (lz, ly, lx) float32."""
        out = np.empty((self.lz, self.ly, self.lx), dtype=np.float32)
        _call("lbm_closure_nut", self._h, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def particles_solid(self):
        s = np.empty(self.N, dtype=np.int32)
        _call("lbm_particles_solid", self._h, s.ctypes.data_as(ctypes.c_void_p))
        return s


class Trainer:
    """Sample store and network of the learned closure on one device (include/lbm.h)."""

    def __init__(self, capacity, params, device=0):
        params = np.ascontiguousarray(params, dtype=np.float32)
        h = ctypes.c_void_p()
        _call("lbm_trainer_create", int(device), int(capacity), _fptr(params), ctypes.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _call("lbm_trainer_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_target(self, mode):
        """Target of later add_filtered calls: "local" (each node's own
        -tau:S/(2 S:S)/nu) or "plane" (the least-squares value of its
        wall-parallel plane)."""
        _call("lbm_trainer_set_target", self._h, {"local": 0, "plane": 1}[mode])

    def add_filtered(self, fine, nu_coarse, s9_coarse, ystar_coarse):
        n = ctypes.c_longlong()
        _call("lbm_trainer_add_filtered", self._h, fine._h, float(nu_coarse), float(s9_coarse), float(ystar_coarse),
              ctypes.byref(n))
        return n.value

    def add_samples(self, features, targets):
        f = np.ascontiguousarray(features, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        _call("lbm_trainer_add_samples", self._h, _fptr(f), _fptr(t), int(t.size))

    def set_plane_targets(self, values, first=0, count=None):
        """Targets of samples [first, first + count) (default: all from first)
        become values[i mod len(values)]: with first at a snapshot start and
        len(values) the coarse lx, each sample gets its plane's value."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        if count is None:
            count = self.count() - first
        _call("lbm_trainer_set_plane_targets", self._h, int(first), int(count), int(v.size), _fptr(v))

    def samples(self):
        n = ctypes.c_longlong()
        _call("lbm_trainer_samples", self._h, ctypes.byref(n), None, None)
        f = np.empty((n.value, MLP_IN), dtype=np.float32)
        t = np.empty(n.value, dtype=np.float32)
        _call("lbm_trainer_samples", self._h, ctypes.byref(n), _fptr(f), _fptr(t))
        return f, t

    def count(self):
        n = ctypes.c_longlong()
        _call("lbm_trainer_samples", self._h, ctypes.byref(n), None, None)
        return n.value

    def grad(self, first, count):
        g = np.empty(MLP_NPARAM, dtype=np.float32)
        loss = ctypes.c_float()
        _call("lbm_trainer_grad", self._h, int(first), int(count), _fptr(g), ctypes.byref(loss))
        return loss.value, g

    def train(self, first, count, iterations, batch, lr):
        loss = ctypes.c_float()
        _call("lbm_trainer_train", self._h, int(first), int(count), int(iterations), int(batch), float(lr),
              ctypes.byref(loss))
        return loss.value

    def predict(self, first, count):
        out = np.empty(count, dtype=np.float32)
        _call("lbm_trainer_predict", self._h, int(first), int(count), _fptr(out))
        return out

    def params(self):
        p = np.empty(MLP_TOTAL, dtype=np.float32)
        _call("lbm_trainer_params", self._h, _fptr(p))
        return p
