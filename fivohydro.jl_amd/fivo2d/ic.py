"""Initial conditions for the 2+1D solver, as primitive fields on the full grid.

  gubser_*            ideal Gubser flow, src/gubser.jl:20-57 (analytic)
  uniform             transversely uniform state, main2D.jl:173-182 (Bjorken)
  hotspot_event       seeded lumpy event: Gaussian hot spots whose centres are
                      drawn from a 2-D Woods-Saxon profile, Gamma(k=1) amplitudes,
                      charm density from the binary-collision density of the same
                      spots (the pattern of examples2d/04_fluctuating_event.jl:66-73)
  grid_csv            a genuinely 2-D IC from an `x,y,T0,alpha0` grid file,
                      initialize_from_grid_csv! (main2D.jl:205-301): optional
                      Gaussian blur of the file grid, bilinear interpolation,
                      vacuum outside the file, radial taper
  radial_profile      the interpolants of a radial `r,T0,alpha0` profile file,
                      load_initial_interpolants (src/io.jl:51-258)
  elliptic_profile    that profile on an area-preserving elliptical radius, the
                      non-axisymmetric IC of gate G6 (test/test_elliptic2d.jl:48-60)

Every function returns arrays over all Ntot cells (ghosts included), flattened in
the library's flat order i = ix*Nytot + iy, ready for Solver.init_from_prims.
"""
import math

import numpy as np

from .model import T_MIN


def _posden(x):
    return np.where(x > 1e-50, x, 1e-50)


def gubser_den(tau, r, q):
    q2 = q * q
    return 1.0 + 2.0 * q2 * (tau * tau + r * r) + (q2 * q2) * ((tau * tau - r * r) ** 2)


def gubser_vr(tau, r, q):
    q2 = q * q
    return (2.0 * q2 * tau * r) / _posden(1.0 + q2 * tau * tau + q2 * r * r)


def gubser_ur(tau, r, q):
    v = np.clip(gubser_vr(tau, r, q), -0.999999999, 0.999999999)
    return v / np.sqrt(1.0 - v * v)


def gubser_temperature(tau, r, q, Tscale):
    qpos = max(abs(q), 1e-50)
    return Tscale * (2.0 * qpos) ** (2 / 3) / (max(tau, 1e-50) ** (1 / 3) *
                                                _posden(gubser_den(tau, r, q)) ** (1 / 3))


def gubser_Tscale_from_center_T(tau0, q, Tc0):
    qpos = max(abs(q), 1e-50)
    fac = max(tau0, 1e-50) ** (1 / 3) * (1.0 + (qpos * tau0) ** 2) ** (2 / 3)
    return max(Tc0, 1e-50) * fac / (2.0 * qpos) ** (2 / 3)


def _mesh(g):
    X, Y = np.meshgrid(g.xC, g.yC, indexing="ij")
    return X, Y


def gubser_ic(g, tau0=1.0, q=1.0, Tc0=0.6, alpha=-20.0):
    """Analytic ideal Gubser state at tau0 (test/test_gubser2d.jl:97-108)."""
    Tscale = gubser_Tscale_from_center_T(tau0, q, Tc0)
    X, Y = _mesh(g)
    r = np.hypot(X, Y)
    T = gubser_temperature(tau0, r, q, Tscale)
    ur = gubser_ur(tau0, r, q)
    with np.errstate(invalid="ignore", divide="ignore"):
        ux = np.where(r > 0, ur * X / np.where(r > 0, r, 1.0), 0.0)
        uy = np.where(r > 0, ur * Y / np.where(r > 0, r, 1.0), 0.0)
    A = np.full_like(T, alpha)
    return dict(T=T.ravel(), alpha=A.ravel(), ux=ux.ravel(), uy=uy.ravel()), Tscale


def uniform(g, T0=0.5, alpha0=-4.2):
    n = g.Ntot
    return dict(T=np.full(n, T0), alpha=np.full(n, alpha0), ux=np.zeros(n), uy=np.zeros(n))


# ---- seeded hot-spot event generator --------------------------------------------
def _woods_saxon_centres(rng, n, R, a):
    """Rejection-sample n points from rho(r) ~ 1/(1+exp((r-R)/a)) on the plane."""
    out = np.empty((0, 2))
    rmax = R + 10 * a
    while out.shape[0] < n:
        m = 4 * (n - out.shape[0]) + 16
        xy = rng.uniform(-rmax, rmax, size=(m, 2))
        r = np.hypot(xy[:, 0], xy[:, 1])
        acc = rng.uniform(0, 1, size=m) < 1.0 / (1.0 + np.exp((r - R) / a))
        out = np.vstack([out, xy[acc & (r <= rmax)]])
    return out[:n]


def hotspot_spots(seed, nspots, R=6.6, a=0.55):
    """The seeded spots of an event: centres from a 2-D Woods-Saxon (R, a) and
    amplitudes A_k ~ Gamma(k=1), from a Philox stream keyed by `seed`."""
    rng = np.random.Generator(np.random.Philox(seed))
    cen = _woods_saxon_centres(rng, nspots, R, a)
    amp = rng.gamma(1.0, 1.0, size=nspots)
    return cen, amp


def hotspot_table(seed, nspots, R=6.6, a=0.55):
    """(x_k, y_k, A_k) rows of the spots, the layout fivo_init_hotspots takes."""
    cen, amp = hotspot_spots(seed, nspots, R, a)
    return np.ascontiguousarray(np.column_stack([cen[:, 0], cen[:, 1], amp]), dtype=np.float64)


def hotspot_event(g, seed, nspots=150, sigma=0.5, Tpeak=0.45, R=6.6, a=0.55, eos=None,
                  alpha_floor=-30.0):
    """Seeded lumpy event on grid `g`.

    Entropy-like density s(x) = sum_k A_k G_sigma(x - x_k) with centres from a
    2-D Woods-Saxon (R, a) and A_k ~ Gamma(k=1). T ~ s^(1/3), scaled so max T =
    Tpeak, floored at T_MIN. The charm fugacity alpha = log(rho_bin / rho_ref)
    follows the binary-collision density rho_bin = sum_{k<l} A_k A_l G(x_k) G(x_l)
    in the product approximation rho_bin ~ s^2, normalised so alpha = 0 at the
    hottest cell and floored at alpha_floor. Dissipative fields start at zero.
    """
    cen, amp = hotspot_spots(seed, nspots, R, a)
    X, Y = _mesh(g)
    s = np.zeros_like(X)
    inv2s2 = 1.0 / (2 * sigma * sigma)
    for (cx, cy), A in zip(cen, amp):
        d2 = (X - cx) ** 2 + (Y - cy) ** 2
        s += A * np.exp(-d2 * inv2s2)
    s = s / s.max()
    T = Tpeak * np.cbrt(s)
    T = np.maximum(T, T_MIN)
    with np.errstate(divide="ignore"):
        alpha = np.where(s > 0, 2.0 * np.log(np.maximum(s, 1e-300)), alpha_floor)
    alpha = np.maximum(alpha, alpha_floor)
    n = g.Ntot
    return dict(T=T.ravel(), alpha=alpha.ravel(), ux=np.zeros(n), uy=np.zeros(n))


# ---- 2-D grid file -------------------------------------------------------------------
def _smoothstep01_5(x):
    """src/utils.jl:27-29: 0 below 0, 1 above 1, x^3 (10 - 15x + 6x^2) between."""
    x = np.asarray(x, dtype=np.float64)
    mid = (x * x * x) * ((10.0 - 15.0 * x) + 6.0 * (x * x))
    return np.where(x <= 0, 0.0, np.where(x >= 1, 1.0, mid))


def read_grid_csv(csv):
    """The file's uniform grid: (xs, ys, T[nx, ny], alpha[nx, ny]) (main2D.jl:241-252).
    Rows are placed by rounding (x - x_1)/dx to the nearest integer, ties to even."""
    raw = np.loadtxt(csv, delimiter=",", skiprows=1, ndmin=2)
    xs = np.unique(raw[:, 0])
    ys = np.unique(raw[:, 1])
    nx, ny = len(xs), len(ys)
    if raw.shape[0] != nx * ny:
        raise ValueError("grid CSV is not a complete uniform grid")
    dx, dy = xs[1] - xs[0], ys[1] - ys[0]
    ix = np.rint((raw[:, 0] - xs[0]) / dx).astype(np.int64)
    iy = np.rint((raw[:, 1] - ys[0]) / dy).astype(np.int64)
    T = np.zeros((nx, ny))
    A = np.zeros((nx, ny))
    T[ix, iy] = raw[:, 2]
    A[ix, iy] = raw[:, 3]
    return xs, ys, T, A


def _blur_axis(A, sig, axis):
    """One pass of the normalised Gaussian blur, edge-clamped: samples outside the
    grid are dropped and the weights renormalised (main2D.jl:259-280), summed in
    kernel order."""
    rad = max(1, int(math.ceil(3 * sig)))
    w = [math.exp(-0.5 * ((k / sig) * (k / sig))) for k in range(-rad, rad + 1)]
    n = A.shape[axis]
    acc = np.zeros_like(A)
    wsum = np.zeros_like(A)
    idx = np.arange(n)
    for kk, k in enumerate(range(-rad, rad + 1)):
        src = idx + k
        ok = (src >= 0) & (src < n)
        sh = np.zeros_like(A)
        if axis == 0:
            sh[ok, :] = A[src[ok], :]
            m = ok[:, None] & np.ones_like(A, dtype=bool)
        else:
            sh[:, ok] = A[:, src[ok]]
            m = ok[None, :] & np.ones_like(A, dtype=bool)
        acc = np.where(m, acc + w[kk] * sh, acc)
        wsum = np.where(m, wsum + w[kk], wsum)
    return acc / wsum


def grid_csv(g, csv, taper_width=1.0, smooth_fm=0.0):
    """Primitive fields of initialize_from_grid_csv! (main2D.jl:238-301) on grid `g`:
    bilinear interpolation of the file's T and alpha (0 on and beyond the file's
    edges), a radial taper over `taper_width` fm inside the file's smaller half
    extent (T -> max(w T, T_MIN), alpha -> w alpha), T >= T_MIN; u = 0."""
    xs, ys, Tg, Ag = read_grid_csv(csv)
    dx, dy = xs[1] - xs[0], ys[1] - ys[0]
    if smooth_fm > 0.0:
        Tg = _blur_axis(_blur_axis(Tg, smooth_fm / dx, 0), smooth_fm / dy, 1)
        Ag = _blur_axis(_blur_axis(Ag, smooth_fm / dx, 0), smooth_fm / dy, 1)
    nx, ny = len(xs), len(ys)
    xlo, xhi, ylo, yhi = xs[0], xs[-1], ys[0], ys[-1]
    rmax_data = min(xhi, yhi)
    tw = max(taper_width, 0.0)
    rt0 = max(rmax_data - tw, 0.0)
    X, Y = _mesh(g)

    def bilin(A):
        inside = (X > xlo) & (X < xhi) & (Y > ylo) & (Y < yhi)
        i = np.clip(np.floor((X - xlo) / dx).astype(np.int64), 0, nx - 2)
        j = np.clip(np.floor((Y - ylo) / dy).astype(np.int64), 0, ny - 2)
        tx = (X - xs[i]) / dx
        ty = (Y - ys[j]) / dy
        v = ((1 - tx) * (1 - ty) * A[i, j] + tx * (1 - ty) * A[i + 1, j]
             + (1 - tx) * ty * A[i, j + 1] + tx * ty * A[i + 1, j + 1])
        return np.where(inside, v, 0.0)

    T = bilin(Tg)
    a = bilin(Ag)
    r = np.hypot(X, Y)
    if tw > 0:
        tap = r > rt0
        w = 1.0 - _smoothstep01_5((r - rt0) / tw)
        T = np.where(tap, np.maximum(w * T, T_MIN), T)
        a = np.where(tap, w * a, a)
    T = np.maximum(T, T_MIN)
    n = g.Ntot
    return dict(T=T.ravel(), alpha=a.ravel(), ux=np.zeros(n), uy=np.zeros(n))


def write_grid_csv(path, xs, ys, T, alpha):
    """Write an `x,y,T0,alpha0` grid file (the format grid_csv reads)."""
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    rows = np.column_stack([X.ravel(), Y.ravel(), np.asarray(T).ravel(), np.asarray(alpha).ravel()])
    np.savetxt(path, rows, delimiter=",", header="x,y,T0,alpha0", comments="", fmt="%.17g")


# ---- radial profile file -----------------------------------------------------------
def radial_profile(csv, fugacity_kind="alpha", taper_width=1.0, interp_kind="linear", dup_r_tol=0.0):
    """load_initial_interpolants (src/io.jl:51-258) for a CSV with columns r, T0 and
    alpha0 (or lambda0), optionally nu_r0 / nu_r (physical, converted to the stored
    sign, which is the same here: DISS_SIGN = +1, src/constants.jl:69) or nur0 / nur
    (stored). Rows are sorted by r (stable), repeated r (within dup_r_tol) keep their
    first row, and every column is interpolated linearly in r. Returns (itpT, itpF,
    itpNur, rmax_data), vectorised functions of r >= 0:
      r <= r_min: the value at r_min (T floored at T_MIN);
      r_min < r <= rmax_data: the interpolant, blended over the last taper_width fm by
        w = 1 - smoothstep01_5((r - r_t0)/taper_width), r_t0 = max(rmax_data - taper_width,
        r_min), towards T_MIN, the vacuum fugacity (alpha 0, lambda 1) and nu 0;
      r > rmax_data: T_MIN, the vacuum fugacity, 0.
    Only interp_kind = "linear" is built (the reference's optional cubic resampling is
    not)."""
    if interp_kind != "linear":
        raise ValueError(f"radial_profile: interp_kind {interp_kind!r} is not supported (only 'linear')")
    if fugacity_kind not in ("alpha", "lambda"):
        raise ValueError("fugacity_kind must be 'alpha' or 'lambda'")
    with open(csv) as fh:
        header = [h.strip() for h in fh.readline().strip().split(",")]
    raw = np.loadtxt(csv, delimiter=",", skiprows=1, ndmin=2)
    col = {h: raw[:, k] for k, h in enumerate(header)}
    fname = "alpha0" if fugacity_kind == "alpha" else "lambda0"
    for need in ("r", "T0", fname):
        if need not in col:
            raise ValueError(f"radial_profile: column {need!r} missing from {csv}")
    r, T0, f0 = col["r"], col["T0"], col[fname]
    nur = None
    for k in ("nu_r0", "nu_r", "nur0", "nur"):
        if k in col:
            nur = col[k].copy()
            break
    if nur is None:
        nur = np.zeros_like(r)
    if len(r) < 2:
        raise ValueError(f"Initial profile CSV must contain at least 2 rows; got {len(r)}.")
    if not (np.all(np.isfinite(r)) and np.all(np.isfinite(T0)) and np.all(np.isfinite(f0)) and np.all(np.isfinite(nur))):
        raise ValueError("Initial profile CSV contains NaN/Inf values.")
    p = np.argsort(r, kind="stable")
    r, T0, f0, nur = r[p], T0[p], f0[p], nur[p]
    keep = np.ones(len(r), dtype=bool)
    tol = max(dup_r_tol, 0.0)
    keep[1:] = np.abs(np.diff(r)) > tol
    r, T0, f0, nur = r[keep], T0[keep], f0[keep], nur[keep]
    if len(r) < 2:
        raise ValueError("After removing duplicate r entries, fewer than 2 points remain.")
    if fugacity_kind == "lambda":
        f0 = np.maximum(f0, 1e-300)
    rmin_data, rmax_data = float(r[0]), float(r[-1])
    T_floor = T_MIN
    f_vac = 0.0 if fugacity_kind == "alpha" else 1.0
    tw = max(taper_width, 0.0)
    r_t0 = max(rmax_data - tw, rmin_data)
    inv_tw = 1.0 / tw if tw > 0 else 0.0

    def lin(v, x):
        j = np.clip(np.searchsorted(r, x, side="right") - 1, 0, len(r) - 2)
        t = (x - r[j]) / (r[j + 1] - r[j])
        return (1.0 - t) * v[j] + t * v[j + 1]

    def make(v, floor, vac):
        v0 = float(v[0])

        def f(rr):
            x = np.asarray(rr, dtype=np.float64)
            inner = lin(v, np.clip(x, rmin_data, rmax_data))
            if tw > 0:
                w = 1.0 - _smoothstep01_5((x - r_t0) * inv_tw)
                tap = w * inner + (1.0 - w) * (vac if floor is None else floor)
                inner = np.where(x >= r_t0, tap, inner)
            out = np.where(x <= rmin_data, v0, np.where(x <= rmax_data, inner, vac if floor is None else floor))
            return np.maximum(out, floor) if floor is not None else out
        return f

    return make(T0, T_floor, None), make(f0, None, f_vac), make(nur, None, 0.0), rmax_data


def elliptic_profile(g, csv, eps2=0.25, taper_width=1.0):
    """Gate G6's IC (test/test_elliptic2d.jl:48-60): the radial profile of `csv`
    evaluated at r = hypot(x/a, y/b), a = sqrt(1 - eps2), b = sqrt(1 + eps2) (area
    preserving), at rest."""
    itpT, itpF, _, _ = radial_profile(csv, taper_width=taper_width)
    a, b = math.sqrt(1 - eps2), math.sqrt(1 + eps2)
    X, Y = _mesh(g)
    r = np.hypot(X / a, Y / b)
    n = g.Ntot
    return dict(T=itpT(r).ravel(), alpha=itpF(r).ravel(), ux=np.zeros(n), uy=np.zeros(n))
