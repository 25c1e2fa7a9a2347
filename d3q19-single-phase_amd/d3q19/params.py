"""Run-time constants of the channel solver, computed as Channel-Flow/para.f90 does.

Every value is formed with the same double-precision operations in the same
order as para.f90:59-174 (Python floats are IEEE doubles), so the numbers
handed to the device library are the reference's bit for bit.
"""
import math

# Velocity set and tables, Channel-Flow/para.f90:201-210.
CIX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0)
CIY = (0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1)
CIZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1)
IPOPP = (0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15)


def para(lx, laminar=False, mrttype=None, visc=None, rstar=None):
    """Physical and MRT constants of Channel-Flow/para.f90:59-174 for lx wall-normal nodes.

    laminar selects the branch of para.f90:59 (turbulent: :60-73, laminar: :74-88).
    mrttype overrides the branch's MRTtype (:70, :84).  visc/rstar override the
    turbulent branch's nu = 0.0036 and Re* = 180 (other Re_tau configurations).
    """
    rho0 = 1.0
    p = {}
    nx = float(lx)
    if not laminar:
        p["visc"] = 0.0036 if visc is None else float(visc)
        p["rstar"] = 180.0 if rstar is None else float(rstar)
        p["ustar"] = 2.0 * p["rstar"] * p["visc"] / nx
        p["force_in_y"] = 2. * rho0 * p["ustar"] * p["ustar"] / nx
        p["ystar"] = p["visc"] / p["ustar"]
        mrt = 1
    else:
        p["rstar"] = 20.0
        p["ustar"] = 0.05
        p["visc"] = 2.0 * p["ustar"] * nx / p["rstar"]
        p["force_in_y"] = 8.0 * p["visc"] * p["ustar"] / (nx * nx)
        p["ystar"] = p["visc"] / p["ustar"]
        mrt = 2
    p["force_mag"] = 1.0
    if mrttype is not None:
        mrt = int(mrttype)
    p["mrttype"] = mrt

    p["tau"] = 3.0 * p["visc"] + 0.5
    s9 = 1.0 / p["tau"]
    p["s9"] = s9
    p["s13"] = s9
    if mrt == 1:
        p.update(s1=1.5, s2=1.4, s4=1.2, s10=1.4, s16=1.98,
                 omegepsl=0.0, omegepslj=-(475.0 / 63.0), omegxx=0.0)
    elif mrt == 2:
        p.update(s1=s9, s2=s9, s4=s9, s10=s9, s16=s9,
                 omegepsl=3.0, omegepslj=-(11.0 / 2.0), omegxx=-(1.0 / 2.0))
    elif mrt == 3:
        p.update(s1=1.8, s2=1.8, s4=s9, s10=1.8, s16=1.8,
                 omegepsl=3.0, omegepslj=-(11.0 / 2.0), omegxx=-(1.0 / 2.0))
    else:
        raise ValueError(f"MRTtype {mrt}: para.f90 defines 1, 2 and 3")

    coef3, coef4 = 8.0, -4.0
    p.update(coef1=-(2.0 / 3.0), coef2=-11.0, coef3=coef3, coef4=coef4, coef5=2.0,
             coef3i=1.0 / coef3, coef4i=1.0 / coef4, val8=12.0,
             val1i=1.0 / 19.0, val2i=1.0 / 2394.0, val3i=1.0 / 252.0, val4i=1.0 / 10.0,
             val5i=1.0 / 40.0, val6i=1.0 / 36.0, val7i=1.0 / 72.0, val8i=1.0 / 12.0,
             val9i=1.0 / 24.0, ww0=1.0 / 3.0, ww1=1.0 / 18.0, ww2=1.0 / 36.0)
    return p


def loglaw_profile(lx, ustar, ystar):
    """Streamwise velocity of initvel's law-of-the-wall start, Channel-Flow/initial.f90:104-115.

    Returns uy(ix) for ix = 1..lx (linear for y+ < 10.8, ln(y+)/0.41 + 5 above),
    mirrored about the channel centre; nxh = lx/2 as for an even lx.
    """
    prof = [0.0] * lx
    for i in range(1, lx // 2 + 1):
        yplus = (float(i) - 0.5) / ystar
        if yplus < 10.8:
            v = yplus * ustar
        else:
            v = (math.log(yplus) / 0.41 + 5.0) * ustar
        prof[i - 1] = v
        prof[lx - i] = v
    return prof
