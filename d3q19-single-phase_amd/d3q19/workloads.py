"""Seeded synthetic initial fields for the channel workloads.

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014) evaluated on a
counter: value k of stream `seed` is splitmix64(seed * 2^32 + k), mapped to a
53-bit uniform double in [0, 1).  The generator is fully specified here so
the same inputs can be produced by any driver (Fortran, C, Python).

Grids use the reference's axes: x wall-normal (lx), y streamwise (ly),
z spanwise (lz); fields are returned as (lz, ly, lx) C-ordered arrays, the
memory order of the reference's (lx, ly, lz) Fortran arrays.
"""
import math

import numpy as np

from .params import loglaw_profile

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """SplitMix64 finaliser applied to x + golden-gamma (vectorised, uint64)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform(seed, count, offset=0):
    """count uniforms in [0, 1) from stream `seed`, starting at counter `offset`."""
    k = np.arange(offset, offset + count, dtype=np.uint64) + np.uint64(seed) * np.uint64(1 << 32)
    return (splitmix64(k) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def laminar_parabola_noise(lx, ly, lz, seed=1, uc=0.05, amp=0.05):
    """u_y = uc (1 - ((i - 0.5 - lx/2)/(lx/2))^2) plus iid U(-amp, amp)*uc on every
    component of every node (streams: ux, uy, uz take counters [0,N), [N,2N), [2N,3N))."""
    n = lx * ly * lz
    h = lx / 2.0
    i = np.arange(1, lx + 1, dtype=np.float64)
    prof = uc * (1.0 - ((i - 0.5 - h) / h) ** 2)
    r = uniform(seed, 3 * n)
    noise = (2.0 * r - 1.0) * amp * uc
    ux = noise[:n].reshape(lz, ly, lx).copy()
    uy = (np.broadcast_to(prof, (lz, ly, lx)) + noise[n:2 * n].reshape(lz, ly, lx)).copy()
    uz = noise[2 * n:].reshape(lz, ly, lx).copy()
    return ux, uy, uz


def _divfree_planes(lx, ly, lz, seed, nmodes, kmax, near_wall):
    """Wall envelope g, its derivative dg (length lx) and the (lz, ly) plane
    fields A..E of the random-phase stream function (see loglaw_divfree)."""
    x = np.arange(1, lx + 1, dtype=np.float64) - 0.5
    y = np.arange(1, ly + 1, dtype=np.float64) - 0.5
    z = np.arange(1, lz + 1, dtype=np.float64) - 0.5
    if near_wall:
        g = np.sin(2.0 * math.pi * x / lx) ** 2
        dg = 2.0 * np.sin(2.0 * math.pi * x / lx) * np.cos(2.0 * math.pi * x / lx) * (2.0 * math.pi / lx)
    else:
        g = np.sin(math.pi * x / lx) ** 2
        dg = 2.0 * np.sin(math.pi * x / lx) * np.cos(math.pi * x / lx) * (math.pi / lx)
    r = uniform(seed, 9 * nmodes)
    # u' separates into wall envelopes times (z, y) plane fields:
    #   u_x = g A,  u_y = g B - g' C,  u_z = g' D - g E
    A = np.zeros((lz, ly))
    B = np.zeros((lz, ly))
    C = np.zeros((lz, ly))
    D = np.zeros((lz, ly))
    E = np.zeros((lz, ly))
    for m in range(nmodes):
        a = 2.0 * r[9 * m: 9 * m + 3] - 1.0
        ky = 2.0 * math.pi * (1 + int(r[9 * m + 3] * kmax[0])) / ly
        kz = 2.0 * math.pi * (1 + int(r[9 * m + 4] * kmax[1])) / lz
        ph = 2.0 * math.pi * r[9 * m + 5: 9 * m + 8]
        arg = ky * y[None, :] + kz * z[:, None]
        c = [np.cos(arg + ph[k]) for k in range(3)]
        s = [np.sin(arg + ph[k]) for k in range(3)]
        # psi_k = a_k g(x) cos(arg + ph_k); u = curl psi
        A += a[2] * (-ky) * s[2] - a[1] * (-kz) * s[1]   # d psi_z/dy - d psi_y/dz
        B += a[0] * (-kz) * s[0]                          # d psi_x/dz
        C += a[2] * c[2]                                  # d psi_z/dx = g' C
        D += a[1] * c[1]                                  # d psi_y/dx = g' D
        E += a[0] * (-ky) * s[0]                          # d psi_x/dy
    return g, dg, A, B, C, D, E


def loglaw_divfree(lx, ly, lz, ustar, ystar, seed=2, amp=0.10, nmodes=8, kmax=(3, 3), near_wall=False):
    """Log-law mean profile (Channel-Flow/initial.f90:104-115) plus a
    divergence-free random-phase perturbation u' = curl(psi).

    psi_c(x, y, z) = sum_m a_{c,m} g(x) cos(ky_m y + kz_m z + phi_{c,m}) with the
    wall envelope g(x) = sin^2(pi x / lx) (x = i - 0.5), or sin^2(2 pi x / lx)
    with near_wall (peaks at a quarter of the width from each wall), integer
    wavenumbers 1..kmax[0] in y and 1..kmax[1] in z on the periodic box;
    derivatives are taken analytically, so div u' = 0 exactly in the
    continuum.  u' is scaled so that max|u'| equals amp times the centreline
    mean velocity.
    """
    prof = np.asarray(loglaw_profile(lx, ustar, ystar))
    g, dg, A, B, C, D, E = _divfree_planes(lx, ly, lz, seed, nmodes, kmax, near_wall)
    ux = A[:, :, None] * g[None, None, :]
    uy = B[:, :, None] * g[None, None, :] - C[:, :, None] * dg[None, None, :]
    uz = D[:, :, None] * dg[None, None, :] - E[:, :, None] * g[None, None, :]
    umax = max(np.abs(ux).max(), np.abs(uy).max(), np.abs(uz).max())
    scale = amp * prof.max() / umax
    ux *= scale
    uy *= scale
    uz *= scale
    uy += prof[None, None, :]
    return ux, uy, uz


def loglaw_divfree_terms(lx, ly, lz, ustar, ystar, seed=2, amp=0.10, nmodes=8, kmax=(3, 3), near_wall=False):
    """The field of loglaw_divfree as separable terms [(component, plane
    (lz, ly), profile (lx)), ...] for Channel.set_velocity_separable, which
    sums them on the device: no 3 x N host arrays.  The perturbation's
    maximum is found plane by plane, wall-normal index by wall-normal index,
    with the same products and differences, so the scale is loglaw_divfree's;
    the device sums the scaled terms, so the fields agree with
    loglaw_divfree's to rounding, not bit for bit (separable_fields gives
    the device's values)."""
    prof = np.asarray(loglaw_profile(lx, ustar, ystar))
    g, dg, A, B, C, D, E = _divfree_planes(lx, ly, lz, seed, nmodes, kmax, near_wall)
    umax = float(np.abs(A).max() * np.abs(g).max())
    b1 = np.empty_like(A)
    b2 = np.empty_like(A)
    for i in range(lx):
        for P, p, Q, q in ((B, g[i], C, dg[i]), (D, dg[i], E, g[i])):
            np.multiply(P, p, out=b1)
            np.multiply(Q, q, out=b2)
            np.subtract(b1, b2, out=b1)
            np.abs(b1, out=b1)
            umax = max(umax, float(b1.max()))
    scale = amp * prof.max() / umax
    one = np.ones((lz, ly))
    return [(0, A * scale, g), (1, one, prof), (1, B * scale, g), (1, -(C * scale), dg),
            (2, D * scale, dg), (2, -(E * scale), g)]


def separable_fields(terms, lx, ly, lz):
    """Host evaluation of separable terms exactly as the device sums them
    (each component from 0.0, term by term in the order given, product then
    add); small grids only."""
    u = [np.zeros((lz, ly, lx)) for _ in range(3)]
    for c, P, Q in terms:
        u[c] = u[c] + np.asarray(P).reshape(lz, ly)[:, :, None] * np.asarray(Q).reshape(lx)[None, None, :]
    return tuple(u)


def separable_at(terms, ix, iy, iz):
    """The separable field's (ux, uy, uz) at nodes (ix, iy, iz) (0-based index
    arrays), summed as the device sums it."""
    ix, iy, iz = (np.asarray(a) for a in (ix, iy, iz))
    u = [np.zeros(ix.shape) for _ in range(3)]
    for c, P, Q in terms:
        u[c] = u[c] + np.asarray(P)[iz, iy] * np.asarray(Q)[ix]
    return tuple(u)


def sphere_positions(lx, ly, lz, n, rad, gap=3.0, gap_w=3.0, seed=3, max_tries=2_000_000):
    """n sphere centres in node coordinates (node i at i - 0.5, 1-based; walls at
    x = 0 and x = lx), uniform by rejection sampling from stream `seed`:
    surface gaps >= gap between spheres (minimum image in y, z) and >= gap_w to
    the walls.  Raises if the spheres cannot be placed."""
    pos = np.empty((n, 3))
    xlo, xhi = rad + gap_w, lx - rad - gap_w
    if xhi <= xlo:
        raise ValueError("channel too narrow for the spheres")
    dmin2 = (2.0 * rad + gap) ** 2
    k = 0
    tries = 0
    batch = 4096
    off = 0
    while k < n:
        r = uniform(seed, 3 * batch, offset=off)
        off += 3 * batch
        for j in range(batch):
            tries += 1
            if tries > max_tries:
                raise ValueError(f"placed only {k} of {n} spheres")
            c = np.array([xlo + (xhi - xlo) * r[3 * j], ly * r[3 * j + 1], lz * r[3 * j + 2]])
            if k:
                d = pos[:k] - c
                d[:, 1] -= ly * np.round(d[:, 1] / ly)
                d[:, 2] -= lz * np.round(d[:, 2] / lz)
                if np.min(np.einsum("ij,ij->i", d, d)) < dmin2:
                    continue
            pos[k] = c
            k += 1
            if k == n:
                break
    return pos
