"""CPU oracle: a NumPy restatement of JathenAX's periodic isothermal stage update.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product path.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to compare against or to time as the baseline.  Nothing in
``jathenax_amd`` imports it, and the product never falls back to it.

Every function follows one reference function, cited as ``file:line`` of the JathenAX
source.  The floating-point expression order (operand grouping, divisions by constants,
reciprocal-then-multiply in recovery) is kept exactly as in the reference, because the
limiters branch on comparisons that flip on one-ulp differences.  NumPy never contracts
``a*b+c`` into an FMA, so in float64 this restatement is the IEEE evaluation of the
reference expressions.  Python-float constants follow NumPy 2's weak promotion (NEP 50),
which matches JAX's weak typing: they are rounded to the state dtype.

Pinning status (what checks this restatement, see tests/test_oracle_*.py):
  * reconstructions, Riemann solvers, viscous fluxes, stage driver, recovery, timestep,
    forcing seed path and kick normalisation: golden vectors written by
    tests/golden/make_reference_vectors.py from the reference functions themselves, plus
    the reference tests' closed forms and invariants, ported;
  * reductions (sum/mean/min/max): NumPy pairwise order, not XLA's; agreement with the
    JAX trajectory is to round-off, not bitwise (the reference cannot run here: JAX is
    not installed), so whole-trajectory bitwise parity with JAX itself is unpinned.
"""

from fractions import Fraction
from math import exp, isfinite, log, pi, sqrt

import numpy as np

# ---------------------------------------------------------------------------------------
# Face reconstructions (jathenax/fluxes.py:6-131).  ``axis`` is the array axis of the
# sweep: 1 = x faces, 0 = y faces.  Each returns (left, right) at each cell's low face:
# left is the high edge of the cell below, right is the low edge of the cell itself.
# ---------------------------------------------------------------------------------------


def _nb(W, offset, axis):
    """Neighbour ``offset`` cells away along ``axis`` with periodic wrap (jnp.roll(W, -offset))."""
    return np.roll(W, -offset, axis=axis)


def dc_face_states(W, axis):
    # fluxes.py:6-8
    return _nb(W, -1, axis), W


def plm_face_states(W, axis):
    # fluxes.py:11-18 (harmonic limiter)
    dl = W - _nb(W, -1, axis)
    dr = _nb(W, 1, axis) - W
    prod = dl * dr
    mono = prod > 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        slope = np.where(mono, prod / np.where(mono, dl + dr, 1.0), 0.0)
    return _nb(W + slope, -1, axis), W - slope


def ppm4_face_states(W, axis):
    # fluxes.py:21-42
    qm2, qm1 = _nb(W, -2, axis), _nb(W, -1, axis)
    qp1, qp2 = _nb(W, 1, axis), _nb(W, 2, axis)
    lo = (7.0 * (W + qm1) - (qm2 + qp1)) / 12.0
    hi = (7.0 * (W + qp1) - (qm1 + qp2)) / 12.0
    lo = np.fmin(np.fmax(lo, np.fmin(W, qm1)), np.fmax(W, qm1))
    hi = np.fmin(np.fmax(hi, np.fmin(W, qp1)), np.fmax(W, qp1))
    dhi = hi - W
    dlo = lo - W
    ext = dhi * dlo >= 0.0
    lo = np.where(np.abs(dlo) >= 2.0 * np.abs(dhi), W - 2.0 * dhi, lo)
    hi = np.where(np.abs(dhi) >= 2.0 * np.abs(dlo), W - 2.0 * dlo, hi)
    lo = np.where(ext, W, lo)
    hi = np.where(ext, W, hi)
    return _nb(hi, -1, axis), lo


def _ppmx_limit_edge(a, b, e, ca, cb):
    # fluxes.py:45-53
    curv = 3.0 * ((a + b) - 2.0 * e)
    same = (((curv > 0) & (ca > 0) & (cb > 0))
            | ((curv < 0) & (ca < 0) & (cb < 0)))
    mag = np.fmin(np.abs(ca), np.abs(cb))
    lim = np.copysign(np.fmin(1.25 * mag, np.abs(curv)), curv)
    corr = 0.5 * (a + b) - np.where(same, lim, 0.0) / 6.0
    return np.where((a - e) * (b - e) > 0.0, corr, e)


def ppmx_face_states(W, axis):
    # fluxes.py:56-84 (smooth-extremum preserving PPM)
    qm2, qm1 = _nb(W, -2, axis), _nb(W, -1, axis)
    qp1, qp2 = _nb(W, 1, axis), _nb(W, 2, axis)
    lo = (7.0 * (W + qm1) - (qm2 + qp1)) / 12.0
    hi = (7.0 * (W + qp1) - (qm1 + qp2)) / 12.0
    d2l = (qm2 + W) - 2.0 * qm1
    d2c = (qm1 + qp1) - 2.0 * W
    d2r = (W + qp2) - 2.0 * qp1
    lo = _ppmx_limit_edge(qm1, W, lo, d2l, d2c)
    hi = _ppmx_limit_edge(W, qp1, hi, d2c, d2r)
    dhi, dlo = hi - W, lo - W
    ext = (dhi * (W - lo) <= 0.0) | ((qm1 - W) * (W - qp1) <= 0.0)
    d2 = 6.0 * (lo + hi - 2.0 * W)
    same = (((d2 > 0) & (d2l > 0) & (d2c > 0) & (d2r > 0))
            | ((d2 < 0) & (d2l < 0) & (d2c < 0) & (d2r < 0)))
    mag = np.fmin(np.abs(d2c), np.fmin(np.abs(d2l), np.abs(d2r)))
    lim = np.copysign(np.fmin(1.25 * mag, np.abs(d2)), d2)
    scale = np.fmax(np.abs(qm1), np.fmax(np.abs(W), np.abs(qp1)))
    use = same & (np.abs(d2) > 1e-12 * scale)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(use, lim / np.where(use, d2, 1.0), 0.0)
    lo_m = np.where(np.abs(dlo) >= 2.0 * np.abs(dhi), W - 2.0 * dhi, lo)
    hi_m = np.where(np.abs(dhi) >= 2.0 * np.abs(dlo), W - 2.0 * dlo, hi)
    lo = np.where(ext, W + dlo * ratio, lo_m)
    hi = np.where(ext, W + dhi * ratio, hi_m)
    return _nb(hi, -1, axis), lo


def _wenoz_high_edge(qm2, qm1, q, qp1, qp2, w0, w1, w2):
    # fluxes.py:87-93
    a0, a1, a2 = 0.1 * w0, 0.6 * w1, 0.3 * w2
    f0 = 2.0 * qm2 - 7.0 * qm1 + 11.0 * q
    f1 = -qm1 + 5.0 * q + 2.0 * qp1
    f2 = 2.0 * q + 5.0 * qp1 - qp2
    return (f0 * a0 + f1 * a1 + f2 * a2) / (6.0 * (a0 + a1 + a2))


def wenoz_epsilon(dtype):
    """WENO-Z epsilon, max(1e-42, smallest normal of dtype) (fluxes.py:108)."""
    return max(1e-42, float(np.finfo(dtype).tiny))


def wenoz_face_states(W, axis):
    # fluxes.py:96-117
    qm2, qm1 = _nb(W, -2, axis), _nb(W, -1, axis)
    qp1, qp2 = _nb(W, 1, axis), _nb(W, 2, axis)
    b0 = (13.0 / 12.0) * (qm2 + W - 2.0 * qm1) ** 2 + 0.25 * (qm2 + 3.0 * W - 4.0 * qm1) ** 2
    b1 = (13.0 / 12.0) * (qm1 + qp1 - 2.0 * W) ** 2 + 0.25 * (qm1 - qp1) ** 2
    b2 = (13.0 / 12.0) * (qp2 + W - 2.0 * qp1) ** 2 + 0.25 * (qp2 + 3.0 * W - 4.0 * qp1) ** 2
    tau = np.abs(b0 - b2)
    eps = wenoz_epsilon(W.dtype)
    d0, d1, d2 = b0 + eps, b1 + eps, b2 + eps
    dmin = np.fmin(d0, np.fmin(d1, d2))
    scale = np.fmax(tau, dmin)
    base = (dmin / scale) ** 2
    contrast = tau / scale
    w0 = base + (contrast * (dmin / d0)) ** 2
    w1 = base + (contrast * (dmin / d1)) ** 2
    w2 = base + (contrast * (dmin / d2)) ** 2
    hi = _wenoz_high_edge(qm2, qm1, W, qp1, qp2, w0, w1, w2)
    lo = _wenoz_high_edge(qp2, qp1, W, qm1, qm2, w2, w1, w0)
    return _nb(hi, -1, axis), lo


RECONSTRUCTIONS = {
    "dc": dc_face_states, "plm": plm_face_states, "ppm4": ppm4_face_states,
    "ppmx": ppmx_face_states, "wenoz": wenoz_face_states,
}


def reconstruct_faces(W, axis, reconstruct, dfloor):
    # fluxes.py:126-131: PPMX and WENO-Z floor the face density.
    left, right = RECONSTRUCTIONS[reconstruct](W, axis)
    if reconstruct in ("ppmx", "wenoz"):
        left = left.copy()
        right = right.copy()
        left[..., 0] = np.fmax(left[..., 0], dfloor)
        right[..., 0] = np.fmax(right[..., 0], dfloor)
    return left, right


# ---------------------------------------------------------------------------------------
# Isothermal Riemann solvers (fluxes.py:134-219).  Arguments are normal/tangent ordered;
# results are (mass, normal momentum, tangential momentum).
# ---------------------------------------------------------------------------------------


def _roe_flux(rl, rr, ul, ur, tl, tr, cs):
    # fluxes.py:134-172
    mnl, mnr = rl * ul, rr * ur
    mtl, mtr = rl * tl, rr * tr
    pl, pr = (cs * cs) * rl, (cs * cs) * rr
    fl = (mnl, mnl * ul + pl, mnl * tl)
    fr = (mnr, mnr * ur + pr, mnr * tr)
    fa = tuple(0.5 * (a + b) for a, b in zip(fl, fr))
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    inv = 1.0 / (sl + sr)
    vn = (sl * ul + sr * ur) * inv
    vt = (sl * tl + sr * tr) * inv
    wm, wp = vn - cs, vn + cs
    drho, dmn, dmt = rr - rl, mnr - mnl, mtr - mtl
    am = drho * (0.5 + 0.5 * vn / cs) - dmn * 0.5 / cs
    ash = drho * (-vt) + dmt
    ap = drho * (0.5 - 0.5 * vn / cs) + dmn * 0.5 / cs
    cm = -0.5 * np.abs(wm) * am
    csh = -0.5 * np.abs(vn) * ash
    cp = -0.5 * np.abs(wp) * ap
    flux = [fa[0] + cm + cp,
            fa[1] + cm * wm + cp * wp,
            fa[2] + cm * vt + csh + cp * vt]
    sup_l = wm >= 0.0
    sup_r = wp <= 0.0
    rstar = rl + am
    llf = (rstar < 0.0) | (rstar + ap < 0.0)
    speed = 0.5 * np.fmax(np.abs(ul) + cs, np.abs(ur) + cs)
    jump = (drho, dmn, dmt)
    out = []
    for c in range(3):
        f = np.where(sup_l, fl[c], flux[c])
        f = np.where(sup_r, fr[c], f)
        out.append(np.where(llf, fa[c] - speed * jump[c], f))
    return tuple(out)


def _llf_flux(rl, rr, ul, ur, tl, tr, cs):
    # fluxes.py:175-183
    mnl, mnr = rl * ul, rr * ur
    mtl, mtr = rl * tl, rr * tr
    fsum = (mnl + mnr, mnl * ul + mnr * ur + cs ** 2 * (rl + rr), mnl * tl + mnr * tr)
    speed = np.fmax(np.abs(ul) + cs, np.abs(ur) + cs)
    jump = (rr - rl, mnr - mnl, mtr - mtl)
    return tuple(0.5 * (fsum[c] - speed * jump[c]) for c in range(3))


def _hlle_flux(rl, rr, ul, ur, tl, tr, cs):
    # fluxes.py:186-201
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    vn = (sl * ul + sr * ur) * (1.0 / (sl + sr))
    al = np.fmin(vn - cs, ul - cs)
    ar = np.fmax(vn + cs, ur + cs)
    bm = np.where(al < 0.0, al, -1e-20).astype(rl.dtype)
    bp = np.where(ar > 0.0, ar, 1e-20).astype(rl.dtype)
    dl, dr = ul - bm, ur - bp
    fl = (rl * dl, rl * ul * dl + cs ** 2 * rl, rl * tl * dl)
    fr = (rr * dr, rr * ur * dr + cs ** 2 * rr, rr * tr * dr)
    wt = 0.5 * (bp + bm) / (bp - bm)
    return tuple(0.5 * (fl[c] + fr[c]) + wt * (fl[c] - fr[c]) for c in range(3))


RIEMANN_SOLVERS = {"llf": _llf_flux, "hlle": _hlle_flux, "roe": _roe_flux}


def isothermal_flux(left, right, cs, axis, rsolver="roe"):
    # fluxes.py:207-219: y faces swap normal/tangent, scalars ride the mass flux upwind.
    n, t = (1, 2) if axis == 1 else (2, 1)
    f = RIEMANN_SOLVERS[rsolver](left[..., 0], right[..., 0], left[..., n], right[..., n],
                                 left[..., t], right[..., t], cs)
    comps = [f[0], f[1], f[2]] if axis == 1 else [f[0], f[2], f[1]]
    if left.shape[-1] > 3:
        mass = f[0][..., None]
        up = np.where(mass >= 0, left[..., 3:], right[..., 3:])
        sc = mass * up
        comps += [sc[..., k] for k in range(sc.shape[-1])]
    return np.stack(comps, axis=-1)


# ---------------------------------------------------------------------------------------
# Viscous fluxes (fluxes.py:227-266), learned-stress face average (269-280), divergence.
# ---------------------------------------------------------------------------------------


def face_velocity_gradients(W, dx, dy, axis):
    # fluxes.py:227-238
    v = W[..., 1:3]
    dn, dtr = (dx, dy) if axis == 1 else (dy, dx)
    normal = (v - np.roll(v, 1, axis=axis)) / dn
    plus = np.roll(v, -1, axis=1 - axis)
    minus = np.roll(v, 1, axis=1 - axis)
    trans = ((plus + np.roll(plus, 1, axis=axis))
             - (minus + np.roll(minus, 1, axis=axis))) / (4.0 * dtr)
    return (normal, trans) if axis == 1 else (trans, normal)


def smagorinsky_viscosity(dv_dx, dv_dy, dx, dy, smagorinsky_cs):
    # fluxes.py:241-245
    shear = dv_dx[..., 1] + dv_dy[..., 0]
    s2 = 2.0 * (dv_dx[..., 0] ** 2 + dv_dy[..., 1] ** 2) + shear ** 2
    return smagorinsky_cs ** 2 * dx * dy * np.sqrt(s2)


def viscous_momentum_flux(W, dx, dy, viscosity, smagorinsky_cs, axis):
    # fluxes.py:248-266
    dv_dx, dv_dy = face_velocity_gradients(W, dx, dy, axis)
    ux, vy = dv_dx[..., 0], dv_dy[..., 1]
    shear = dv_dx[..., 1] + dv_dy[..., 0]
    if axis == 1:
        normal = (4.0 / 3.0) * ux - (2.0 / 3.0) * vy
        stress = np.stack((normal, shear), axis=-1)
    else:
        normal = (4.0 / 3.0) * vy - (2.0 / 3.0) * ux
        stress = np.stack((shear, normal), axis=-1)
    nu = viscosity
    if smagorinsky_cs > 0.0:
        nu = nu + smagorinsky_viscosity(dv_dx, dv_dy, dx, dy, smagorinsky_cs)
    rho = W[..., 0]
    mu = 0.5 * nu * (rho + np.roll(rho, 1, axis=axis))
    return -mu[..., None] * stress


def stress_face_fluxes(tau):
    # fluxes.py:269-280
    if tau.ndim != 3 or tau.shape[-1] != 3:
        raise ValueError("Expected stress with shape (ny, nx, 3)")
    tx, ty = tau[..., :2], tau[..., 1:]
    return 0.5 * (np.roll(tx, 1, axis=1) + tx), 0.5 * (np.roll(ty, 1, axis=0) + ty)


def flux_divergence(flux_x, flux_y, dx, dy):
    # fluxes.py:283-287
    return ((np.roll(flux_x, -1, axis=1) - flux_x) / dx
            + (np.roll(flux_y, -1, axis=0) - flux_y) / dy)


# ---------------------------------------------------------------------------------------
# RK stages, recovery, timestep, step driver (jathenax/hydro.py).
# ---------------------------------------------------------------------------------------

RK_STAGES = {
    # hydro.py:20-36, rows (gam0, gam1, beta, delta)
    "rk2": ((0.0, 1.0, 1.0, None), (0.5, 0.5, 0.5, None)),
    "rk3": ((0.0, 1.0, 1.0, None), (0.25, 0.75, 0.25, None),
            (2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, None)),
    "rk4": ((0.0, 1.0, 1.193743905974738, None),
            (0.121098479554482, 0.721781678111411, 0.099279895495783, 0.217683334308543),
            (-3.843833699660025, 2.121209265338722, 1.131678018054042, 1.065841341361089),
            (0.546370891121863, 0.198653035682705, 0.310665766509336, 0.0)),
}


def rk_stage(U, U_saved, div_flux, dt, coefficients):
    # hydro.py:39-51 (blended form when gam0 != 0 and gam0 + gam1 == 1 in float64)
    g0, g1, beta, delta = coefficients
    if delta is not None:
        U_saved = U_saved + delta * U
    sdt = beta * dt
    if g0 != 0.0 and g0 + g1 == 1.0:
        return U_saved + (g0 * (U - U_saved) - sdt * div_flux), U_saved
    return g0 * U + g1 * U_saved - sdt * div_flux, U_saved


def recover_isothermal(U, dfloor):
    # hydro.py:54-60: floor written back into U, then reciprocal-times-momentum.
    rho = np.where(U[..., 0] < dfloor, dfloor, U[..., 0]).astype(U.dtype)
    U = np.concatenate((rho[..., None], U[..., 1:]), axis=-1)
    W = np.concatenate((rho[..., None], (1.0 / rho)[..., None] * U[..., 1:]), axis=-1)
    return U, W


def isothermal_timestep(W, dx, dy, cs, cfl, previous_dt, time, tlim, *,
                        viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0):
    # hydro.py:63-89
    lim = np.minimum(np.min(dx / (np.abs(W[..., 1]) + cs)),
                     np.min(dy / (np.abs(W[..., 2]) + cs)))
    if smagorinsky_cs > 0.0:
        nu_max = 0.0
        for axis in (0, 1):
            gx, gy = face_velocity_gradients(W, dx, dy, axis)
            nu_max = np.maximum(nu_max, np.max(smagorinsky_viscosity(gx, gy, dx, dy,
                                                                     smagorinsky_cs)))
        h = min(dx, dy)
        rate = (viscosity + 2.0 * nu_max) / (h * h)
        with np.errstate(divide="ignore"):
            lim = np.minimum(lim, 0.1875 / rate)
    elif viscosity > 0.0:
        h = min(dx, dy)
        lim = np.minimum(lim, 0.25 * (h * h) / viscosity)
    if drag_rate > 0.0:
        lim = np.minimum(lim, 1.0 / drag_rate)
    dt = np.minimum(2.0 * previous_dt, cfl * lim)
    return np.where((time < tlim) & (time + dt > tlim), tlim - time, dt)[()]


def isothermal_step(U, dt, dx, dy, cs, dfloor, integrator, *,
                    viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0,
                    raw_force=None, dedt=0.0, trace=False, reconstruct="ppm4", rsolver="roe",
                    predict_stress=None, sgs_weights=None):
    # hydro.py:92-146
    U = np.asarray(U)
    U, W = recover_isothermal(U, dfloor)
    U_saved = U
    stages = []
    sched = RK_STAGES[integrator]
    for s, coef in enumerate(sched, 1):
        lx, rx = reconstruct_faces(W, 1, reconstruct, dfloor)
        ly, ry = reconstruct_faces(W, 0, reconstruct, dfloor)
        fx = isothermal_flux(lx, rx, cs, 1, rsolver)
        fy = isothermal_flux(ly, ry, cs, 0, rsolver)
        if viscosity > 0.0 or smagorinsky_cs > 0.0:
            fx = fx.copy()
            fy = fy.copy()
            fx[..., 1:3] = fx[..., 1:3] + viscous_momentum_flux(W, dx, dy, viscosity,
                                                                 smagorinsky_cs, 1)
            fy[..., 1:3] = fy[..., 1:3] + viscous_momentum_flux(W, dx, dy, viscosity,
                                                                 smagorinsky_cs, 0)
        if predict_stress is not None:
            prim = W[..., :3]
            tau = np.asarray(predict_stress(sgs_weights, prim))
            if tau.shape != prim.shape or not np.issubdtype(tau.dtype, np.floating):
                raise ValueError(
                    "predict_stress must return a floating array of shape (ny, nx, 3)")
            tx, ty = stress_face_fluxes(tau.astype(U.dtype))
            fx = fx.copy()
            fy = fy.copy()
            fx[..., 1:3] = fx[..., 1:3] + tx
            fy[..., 1:3] = fy[..., 1:3] + ty
        div = flux_divergence(fx, fy, dx, dy)
        U, U_saved = rk_stage(U, U_saved, div, dt, coef)
        if drag_rate > 0.0:
            sdt = coef[2] * dt
            drag = (sdt * drag_rate * W[..., 0, None]) * W[..., 1:3]
            U = U.copy()
            U[..., 1:3] = U[..., 1:3] + (-drag)
        kick = raw_force is not None and s == len(sched)
        if kick:
            before = U
            acc = normalize_force(U, raw_force, dt, dedt)
            U = U.copy()
            U[..., 1:3] = U[..., 1:3] + U[..., 0, None] * (acc * dt)
        if trace:
            stages.append(dict(primitives=W, left_x=lx, right_x=rx, left_y=ly, right_y=ry,
                               flux_x=fx, flux_y=fy, saved=U_saved, conserved=U))
            if kick:
                stages[-1].update(before_kick=before, acceleration=acc)
        U, W = recover_isothermal(U, dfloor)
    if trace:
        return U, tuple(stages)
    return U


# ---------------------------------------------------------------------------------------
# Turbulence driving (jathenax/forcing.py).
# ---------------------------------------------------------------------------------------


def ran_st(seed):
    """Shuffled two-stream L'Ecuyer uniform generator (forcing.py:11-32)."""
    m1, m2 = 2147483563, 2147483399
    s1 = s2 = max(1, seed)
    table = [0] * 32
    for idx in range(39, -1, -1):
        s1 = (40014 * s1) % m1
        if idx < 32:
            table[idx] = s1
    iy = table[0]
    div = 1 + (m1 - 1) // 32
    top = 1.0 - np.finfo(float).eps
    while True:
        s1 = (40014 * s1) % m1
        s2 = (40692 * s2) % m2
        j = iy // div
        iy = table[j] - s2
        table[j] = s1
        if iy < 1:
            iy += m1 - 1
        yield min((1.0 / m1) * iy, top)


def ran_gaussian_st(seed):
    """Polar Box-Muller, second value of each pair first (forcing.py:35-45)."""
    u = ran_st(seed)
    while True:
        v1 = 2.0 * next(u) - 1.0
        v2 = 2.0 * next(u) - 1.0
        r2 = v1 * v1 + v2 * v2
        if 0.0 < r2 < 1.0:
            f = sqrt(-2.0 * log(r2) / r2)
            yield v2 * f
            yield v1 * f


def cartesian_modes(nlow, nhigh, kx_bounds, ky_bounds):
    # forcing.py:48-60, kx outer, ky inner
    if not 1 <= nlow < nhigh:
        raise ValueError("parabolic driving requires 1 <= nlow < nhigh")
    out = [(kx, ky) for kx in range(kx_bounds[0], kx_bounds[1] + 1)
           for ky in range(ky_bounds[0], ky_bounds[1] + 1)
           if nlow ** 2 <= kx ** 2 + ky ** 2 <= nhigh ** 2]
    if not out:
        raise ValueError("no Cartesian forcing modes in the requested band")
    return np.asarray(out)


def modal_coefficients(modes, lengths, nlow, nhigh, npeak, sol_fraction, rseed,
                       tcorr, dt_update, updates):
    # forcing.py:63-100; the spectrum uses a correctly rounded fma(prefactor, o^2, 1).
    if not nlow <= npeak <= nhigh or not 0.0 <= sol_fraction <= 1.0:
        raise ValueError("npeak must be in the forcing band and sol_fraction in [0, 1]")
    if not np.isfinite(tcorr) or tcorr < 0.0 or not 0.0 < dt_update < np.inf:
        raise ValueError("tcorr must be finite and nonnegative and dt_update positive")
    Lx = lengths[0]
    dk = 2.0 * pi / np.asarray(lengths)
    k = modes * dk
    kiso = np.sqrt(k[:, 0] ** 2 + k[:, 1] ** 2)
    kpeak = npeak * 2.0 * pi / Lx
    band = nhigh * max(dk) - nlow * min(dk)
    pre = Fraction(float(-4.0 / band ** 2))
    power = np.asarray([float(pre * Fraction(float(o)) + 1) for o in (kiso - kpeak) ** 2])
    norm = np.sqrt(np.abs(power))
    norm *= np.sqrt(kpeak / kiso)
    corr = exp(-dt_update / tcorr) if tcorr > 1e-6 else 0.0
    nw = sqrt(1.0 - corr * corr)
    g = ran_gaussian_st(rseed)
    amp = np.empty((updates, len(modes), 2, 2))
    for u in range(updates):
        noise = np.fromiter(g, dtype=float, count=4 * len(modes)).reshape(-1, 2, 2)
        noise *= norm[:, None, None]
        kdn = k[:, 0, None] * noise[:, 0] + k[:, 1, None] * noise[:, 1]
        lon = k[:, :, None] * kdn[:, None] / kiso[:, None, None] ** 2
        noise = sol_fraction * (noise - lon) + (1.0 - sol_fraction) * lon
        amp[u] = noise if u == 0 else corr * amp[u - 1] + nw * noise
    return amp


def forcing_basis(modes, x, y, lengths):
    # forcing.py:103-108
    k = modes * (2.0 * pi / np.asarray(lengths, dtype=x.dtype))
    px = k[:, 0, None] * x
    py = k[:, 1, None] * y
    return np.cos(px), np.sin(px), np.cos(py), np.sin(py)


def render_force(amplitudes, basis):
    # forcing.py:111-122, modes accumulated in order
    cx, sx, cy, sy = basis
    f = np.zeros((cy.shape[1], cx.shape[1], 2), amplitudes.dtype)
    for n in range(len(amplitudes)):
        re = cx[n] * cy[n, :, None] - sx[n] * sy[n, :, None]
        im = cx[n] * sy[n, :, None] + sx[n] * cy[n, :, None]
        f = f + (re[..., None] * amplitudes[n, :, 0] - im[..., None] * amplitudes[n, :, 1])
    return f


def update_force(time, update, raw_force, coefficients, basis, dt_update, tdriv_start):
    # forcing.py:125-134; out-of-table rows render as NaN (mode="fill")
    req = int(np.floor((time - tdriv_start) / dt_update).astype(np.int32))
    req = req if time >= tdriv_start else -1
    if req > update:
        if 0 <= req < len(coefficients):
            row = coefficients[req]
        else:
            row = np.full(coefficients.shape[1:], np.nan, coefficients.dtype)
        raw_force = render_force(row, basis)
    return req, raw_force


def normalize_force(U, raw_force, dt, dedt):
    # forcing.py:137-149
    rho = U[..., 0, None]
    mom = U[..., 1:3]
    mean = np.sum(rho * raw_force, axis=(0, 1)) / np.sum(rho)
    f = raw_force - mean
    fsq = f[..., 0] ** 2 + f[..., 1] ** 2
    m0 = np.mean(0.5 * rho[..., 0] * fsq * dt)
    m1 = np.mean(mom[..., 0] * f[..., 0] + mom[..., 1] * f[..., 1])
    root = np.sqrt(m1 * m1 + 4.0 * m0 * dedt)
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = np.where(m1 >= 0.0, 2.0 * dedt / (m1 + root), (root - m1) / (2.0 * m0))
    scale = np.where(m0 > 1e-30, scale, np.nan)
    return np.where(dedt == 0.0, 0.0, scale * f).astype(U.dtype)


# ---------------------------------------------------------------------------------------
# Time loop (jathenax/run.py:20-69), as a host loop.
# ---------------------------------------------------------------------------------------


def evolve_isothermal(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, *,
                      viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0,
                      forcing=None, dt_update=1.0, tdriv_start=0.0, dedt=0.0,
                      reconstruct="ppm4", rsolver="roe", predict_stress=None, sgs_weights=None,
                      time_dtype=None):
    """Returns (U, time, dt, cycle).  ``time_dtype`` defaults to U.dtype as in the
    reference (run.py:63-64); float64 gives the FP64 time accumulator variant."""
    for name, value in (("viscosity", viscosity), ("smagorinsky_cs", smagorinsky_cs),
                        ("drag_rate", drag_rate), ("dedt", dedt)):
        if not isfinite(value) or value < 0.0:
            raise ValueError(f"{name} must be finite and nonnegative")
    if forcing is not None and (not 0.0 < dt_update < np.inf or not isfinite(tdriv_start)):
        raise ValueError("dt_update must be finite and positive and tdriv_start finite")
    U = np.asarray(U)
    tdt = np.dtype(time_dtype or U.dtype).type
    time, prev, cycle, update = tdt(0.0), tdt(np.inf), 0, -1
    raw = np.zeros_like(U[..., 1:3]) if forcing is not None else None
    while time < tlim and cycle < nlim:
        _, W = recover_isothermal(U, dfloor)
        dt = tdt(isothermal_timestep(W, dx, dy, cs, cfl, prev, time, tlim,
                                     viscosity=viscosity, smagorinsky_cs=smagorinsky_cs,
                                     drag_rate=drag_rate))
        power = 0.0
        if forcing is not None:
            coefficients, basis = forcing
            update, raw = update_force(time, update, raw, coefficients, basis, dt_update,
                                       tdriv_start)
            power = dedt if update >= 0 else 0.0
        U = isothermal_step(U, U.dtype.type(dt), dx, dy, cs, dfloor, integrator,
                            viscosity=viscosity, smagorinsky_cs=smagorinsky_cs,
                            drag_rate=drag_rate, raw_force=raw, dedt=power,
                            reconstruct=reconstruct, rsolver=rsolver,
                            predict_stress=predict_stress, sgs_weights=sgs_weights)
        time, prev, cycle = tdt(time + dt), dt, cycle + 1
    return U, time, prev, cycle


def conserved_sums(U):
    """Box sums in float64 (run.py:186-187 without the dx*dy factor)."""
    return np.asarray(U).sum(axis=(0, 1), dtype=np.float64)
